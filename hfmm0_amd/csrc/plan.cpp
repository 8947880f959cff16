// Plan creation (tree, lists, operator tables, device buffers), the evaluate sequence
// and the C ABI declared in include/hfmm.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <tuple>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/hfmm.h"
#include "common.h"
#include "kernels.h"

namespace hfmm {

void fail(int code, const std::string& msg) { throw Error{code, msg}; }

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

#define HIPCHK(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) fail(HFMM_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kDirV = 0;         // 343 same-level displacements
constexpr int kDirM2M = 343;     // 8 child octants, t = c_parent - c_child
constexpr int kDirL2L = 351;     // 8 child octants, t = c_child - c_parent
constexpr int kNumDirs = 359;

inline bool real_charges(int q_type) { return q_type == HFMM_F64 || q_type == HFMM_F32; }
inline size_t charge_bytes(int q_type) {
    return q_type == HFMM_C128 ? 16 : (q_type == HFMM_F32 ? 4 : 8);
}

struct M2LBlockLevel {
    bool use = false;
    bool fp32 = false;             // kernel precision (M2L of a mixed-precision level)
    int tb = 0;                    // target boxes per workgroup
    int grp = 1;                   // directions per group
    int ps = 0, pt = 0;            // source and target table degrees
    int src_level = 0, tgt_level = 0;
    int n_items = 0;
    int4* items = nullptr;
    int* targets = nullptr;
    unsigned short* dir = nullptr;
    int* off = nullptr;
    int* units1 = nullptr;
    int n_units1 = 0;
    int* units2 = nullptr;
    int n_units2 = 0;
    int* units3 = nullptr;
    int n_units3 = 0;
    void* coax = nullptr;          // A fragments of the coaxial tables, kernel precision
    int* coax_off = nullptr;       // per order m
    int* coax_offq = nullptr;      // per order m: its 4 x 4 fragment (-1: block larger than 4 x 4)
    int coax_stride = 0;
    bool split = false;            // M2M split over items: the targets are cleared first
    int n_zero = 0;                // targets to clear
    int* zero = nullptr;
};

struct XlateLaunch {
    int n_targets = 0;
    int* targets = nullptr;
    int* ptr = nullptr;
    int* src = nullptr;
    unsigned short* dir = nullptr;
    double2* coax = nullptr;
    int64_t coax_stride = 0;
    int ps = 0, pt = 0;            // table degrees of the source / target level
    int src_level = 0, tgt_level = 0;
};

}  // namespace hfmm

using namespace hfmm;

struct hfmm_plan {
    Tree tree;
    Lists lists;
    int64_t n = 0;
    double k = 0.0;                 // |k|
    int conj = 0;                   // k < 0: conjugate charges in and results out
    int precision = HFMM_PREC_FP64;
    int use_graph = 1;
    int depth = 0;
    int pmax = 0;
    int p_level[kMaxLevels] = {0};  // table degree per level (tables, storage, LDS)
    int p_eff[kMaxLevels] = {0};    // degree evaluated per level (<= p_level)
    double h[kMaxLevels] = {0}, R[kMaxLevels] = {0};
    std::vector<int64_t> exp_off;
    int64_t n_coef = 0;
    int64_t n_p2p_pairs = 0;
    int n_leaves = 0, n_far_leaves = 0, p2m_chunk = 8, p_leafmax = 2;
    int n_p2p_fp32 = 0;             // leading leaves of the FP32 near field (phase in FP32)
    bool p2p_lowk = false;          // every near pair has k r <= kP2PLowKr (mixed form)
    std::vector<double> box_maxabs; // setup only: largest |x| of the points of each box
    double* d_floor2 = nullptr;     // per leaf (leaf_box order): coincidence floor on r^2
    unsigned degree_version = 0;    // bumped by set_degrees (graphs that captured this plan are stale)

    std::vector<void*> allocs;
    double *dX = nullptr, *dY = nullptr, *dZ = nullptr;
    int* d_perm = nullptr;
    double* d_centre = nullptr;
    int *d_level = nullptr, *d_pstart = nullptr, *d_count = nullptr;
    int64_t* d_expoff = nullptr;
    int* d_plevel = nullptr;
    double* d_R = nullptr;
    int* d_far_leaves = nullptr;
    LeafDesc* d_leafdesc = nullptr;   // per far leaf (d_far_leaves order): P2M / L2P descriptors
    int *d_leaf_box = nullptr, *d_rng_ptr = nullptr, *d_rng_start = nullptr, *d_rng_pref = nullptr,
        *d_leaf_src_total = nullptr;
    double* d_rot = nullptr;
    double2* d_phase = nullptr;
    DispInfo* d_dinfo = nullptr;
    XlateLaunch m2l[kMaxLevels], m2m[kMaxLevels], l2l[kMaxLevels];
    M2LBlockLevel m2lb[kMaxLevels];
    bool m2l_fp32[kMaxLevels] = {};  // M2L of the level in FP32 (mixed precision)
    M2LBlockLevel m2mb[kMaxLevels], l2lb[kMaxLevels];   // M2M / L2L on matrix cores
    int m2l_kernel = 0;
    int p2p_tile = 2;               // targets per lane of the mixed near field
    double p2p_coord32_tol = 0.0;   // estimated error up to which a leaf takes FP32 coordinates (<= 0: none)
    unsigned char* d_coord32 = nullptr;   // per leaf: FP32 coordinate differences (mixed, FP32 phase)
    int n_leaves_coord32 = 0;
    size_t max_lds = 65536;
    // rotation blocks as MFMA A fragments (E^T for the forward, E for the back rotation)
    double *d_rf1 = nullptr, *d_rf3 = nullptr;
    float *d_rf1_32 = nullptr, *d_rf3_32 = nullptr;
    int *d_rf_off1 = nullptr, *d_rf_off3 = nullptr;
    int* d_rf_off1q = nullptr;   // forward rotation blocks of at most 4 rows: their 4 x 4 fragments
    int* d_rf_off3p = nullptr;   // packed stage-3 units: [8 n1 + n2]
    int rf_stride1 = 0, rf_stride3 = 0;
    float2* d_phase32 = nullptr;
    float2* dM32 = nullptr;
    float2* dL32 = nullptr;         // FP32 copy of the locals that FP32 L2L translations read
    double2 *dM = nullptr, *dL = nullptr, *d_q = nullptr, *d_far = nullptr;
    void* d_qin = nullptr;          // staging for host input
    double2* d_out = nullptr;       // staging for host output

    hipStream_t stream = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    const void* g_q = nullptr;
    void* g_out = nullptr;
    int g_qtype = -1;
    int g_outtype = -1;
    float2* d_out32 = nullptr;      // staging of complex64 output for host pointers
    double2* d_far_e = nullptr;     // far field E (sorted order), allocated at the first field call
    double2* d_field_out = nullptr; // staging of the field for host pointers

    template <typename T>
    T* alloc(size_t count) {
        if (count == 0) count = 1;
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) fail(HFMM_ERR_MEMORY, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        allocs.push_back(p);
        return (T*)p;
    }
    template <typename T>
    T* upload(const std::vector<T>& v) {
        T* p = alloc<T>(v.size());
        if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return p;
    }
    ~hfmm_plan() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    DevBoxes boxes() const { return DevBoxes{d_centre, d_level, d_pstart, d_count, d_expoff}; }
    DevLevels levels() const { return DevLevels{d_plevel, d_R}; }
};

namespace {

void check_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) fail(HFMM_ERR_DEVICE, "no HIP device available");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos)
        fail(HFMM_ERR_DEVICE, std::string("libhfmm is built for gfx950, found ") + prop.gcnArchName);
}

int octant(uint64_t code) { return (int)(code & 7); }

// Units of work of a translation from degree ps to degree pt: forward rotations of the u
// and v blocks of degrees n <= ps, coaxial orders |m| <= min, back rotations (n, row tile
// of the orders 0..n) with n <= pt.  The kernel
// deals list entry u to wave u % 4; the lists are laid out in rows of 4 so that every
// wave gets about the same number of MFMAs: units sorted by cost, each row's units given,
// largest first, to the waves with the least work so far.  Rows are padded with code 255,
// which every stage skips (a degree or order above any table).
void make_units(int ps, int pt, std::vector<int>& units1, std::vector<int>& units2, std::vector<int>& units3,
                bool pack3 = false) {
    std::vector<std::pair<int, int>> u1, u2, u3;   // (cost, code)
    for (int n = 0; n <= ps; ++n) {
        u1.push_back({m2l_tiles(n + 1) * ((n + 4) / 4), n});                   // u block
        if (n > 0) u1.push_back({m2l_tiles(n) * ((n + 3) / 4), n | (1 << 8)});  // v block
    }
    int n3 = 0;
    if (pack3) {   // degrees n1 + n2 = q share a unit (rows 0..7: n1, rows 8..15: n2)
        const int q = std::min(pt, kM2LPackMax);
        for (int n1 = 0; n1 <= q - n1; ++n1) {
            const int n2 = q - n1;
            if (n1 < n2)
                u3.push_back({(n1 + n2 + 5) / 4 + (n1 + n2 + 3) / 4 + 2, 256 | (n1 << 4) | n2});
            else
                u3.push_back({(n1 + 4) / 4 + (n1 + 3) / 4 + 2, n1});
        }
        n3 = q + 1;
    }
    for (int n = n3; n <= pt; ++n)
        for (int mt = 0; mt < m2l_tiles(n + 1); ++mt)
            u3.push_back({(n + 4) / 4 + (n + 3) / 4 + 2, n | (mt << 8)});   // + phase epilogue
    for (int m = 0; m <= std::min(ps, pt); ++m) {
        const int lens = ps + 1 - m, lent = pt + 1 - m;
        u2.push_back({(m ? 2 : 1) * m2l_tiles(lent) * ((lens + 3) / 4), m});
    }
    auto deal = [](std::vector<std::pair<int, int>> u, std::vector<int>& out) {
        std::stable_sort(u.begin(), u.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
            return a.first > b.first;
        });
        while (u.size() % 4) u.push_back({0, 255});
        out.assign(u.size(), 255);
        long load[4] = {0, 0, 0, 0};
        for (size_t r = 0; r < u.size(); r += 4) {
            int order[4] = {0, 1, 2, 3};
            std::stable_sort(order, order + 4, [&](int a, int b) { return load[a] < load[b]; });
            for (int i = 0; i < 4; ++i) {
                out[r + order[i]] = u[r + i].second;
                load[order[i]] += u[r + i].first;
            }
        }
    };
    deal(u1, units1);
    deal(u2, units2);
    deal(u3, units3);
}

// Coaxial tables as A fragments: per distance, per order m = 0..min(ps, pt) the packed
// (pt+1-m) x (ps+1-m) block T[i][k] (row i: target degree m+i, column k: source degree
// m+k), padded to 16x16 tiles; per (row tile, k group, lane) eight values: Re then Im
// of j = 0..3.  coff[m] is the offset of order m within one distance.  Blocks of at most
// kM2LQuadMaxRows squared also get fragments for the 4x4x1 MFMA with its A broadcast from lanes 0..3: per
// 4-row tile t and k group g (t-major), lane l holds Re then Im of T[4t + l % 4][4g + j],
// j = 0..3; they follow the 16x16 fragments of the distance, at coffq[m] (-1 for larger
// blocks).
int coax_fragments(const std::vector<double>& tabs, int n_dist, int ps, int pt, std::vector<double>& cf,
                   std::vector<int>& coff, std::vector<int>& coffq) {
    const int mm = std::min(ps, pt);
    coff.assign(mm + 2, 0);
    for (int m = 0; m <= mm; ++m) coff[m + 1] = coff[m] + m2l_tiles(pt + 1 - m) * m2l_tiles(ps + 1 - m) * 64 * 8;
    int stride = coff[mm + 1];
    coffq.assign(mm + 1, -1);
    for (int m = 0; m <= mm; ++m)
        if (pt + 1 - m <= kM2LQuadMaxRows && ps + 1 - m <= kM2LQuadMaxRows) {
            coffq[m] = stride;
            stride += ((pt + 4 - m) / 4) * ((ps + 4 - m) / 4) * 64 * 8;
        }
    cf.assign((size_t)n_dist * stride, 0.0);
    const int csz = coax_size(pt, ps);
    for (int di = 0; di < n_dist; ++di) {
        const double* tab = &tabs[(size_t)di * csz * 2];
        double* dst = &cf[(size_t)di * stride];
        int off = 0;
        for (int m = 0; m <= mm; ++m) {
            const int lent = pt + 1 - m, lens = ps + 1 - m;
            const int nr = m2l_tiles(lent), nk = m2l_tiles(lens);
            for (int rt = 0; rt < nr; ++rt)
                for (int kg = 0; kg < nk; ++kg)
                    for (int lane = 0; lane < 64; ++lane)
                        for (int j = 0; j < 4; ++j) {
                            const int i = 16 * rt + (lane & 15), k = 16 * kg + 4 * j + (lane >> 4);
                            if (i >= lent || k >= lens) continue;
                            double* f = dst + coff[m] + ((size_t)(rt * nk + kg) * 64 + lane) * 8;
                            f[j] = tab[2 * (off + i * lens + k)];
                            f[4 + j] = tab[2 * (off + i * lens + k) + 1];
                        }
            if (coffq[m] >= 0) {
                const int nt4 = (lent + 3) / 4, nk4 = (lens + 3) / 4;
                for (int t4 = 0; t4 < nt4; ++t4)
                    for (int g4 = 0; g4 < nk4; ++g4)
                        for (int lane = 0; lane < 64; ++lane)
                            for (int j = 0; j < 4; ++j) {
                                const int i = 4 * t4 + (lane & 3), k = 4 * g4 + j;
                                if (i >= lent || k >= lens) continue;
                                double* f = dst + coffq[m] + ((size_t)(t4 * nk4 + g4) * 64 + lane) * 8;
                                f[j] = tab[2 * (off + i * lens + k)];
                                f[4 + j] = tab[2 * (off + i * lens + k) + 1];
                            }
            }
            off += lent * lens;
        }
    }
    return stride;
}

// Work items of target blocks: blocks in space order (the parent region of their first
// target, then octant), so that the kernel's XCD-aware item mapping gives each XCD a
// compact region whose source expansions stay in its L2; entry lists split when there
// are few blocks (if allowed); padded to a multiple of the XCD count with empty items.
std::vector<int4> make_items(const Tree& t, const std::vector<int>& blk_targets, const std::vector<int>& blk_ptr,
                             int tb, bool allow_split) {
    const int nblk = (int)blk_ptr.size() - 1;
    std::vector<int> border(nblk);
    for (int b = 0; b < nblk; ++b) border[b] = b;
    auto region = [&](int b) {
        const int t0 = blk_targets[(size_t)b * tb];
        return t0 < 0 ? ~0ull : t.code[t0];
    };
    std::stable_sort(border.begin(), border.end(), [&](int x, int y) {
        const uint64_t rx = region(x) >> 3, ry = region(y) >> 3;
        return rx != ry ? rx < ry : (region(x) & 7) < (region(y) & 7);
    });
    // measured at N = 1M, k = 40: 8 blocks (level 2) run best split 64 ways (0.12 ms against
    // 0.20 at 16), 64 blocks (level 3) at 16 ways (0.19 ms against 0.25 at 32)
    const int cap = nblk <= 16 ? 64 : 16;
    const int nsplit = allow_split ? std::max(1, std::min(cap, (2048 + nblk - 1) / std::max(nblk, 1))) : 1;
    std::vector<int4> items;
    for (int b : border) {
        const int e0 = blk_ptr[b], e1 = blk_ptr[b + 1];
        const int chunk = std::max(1, (e1 - e0 + nsplit - 1) / nsplit);
        for (int s = e0; s < e1; s += chunk)
            items.push_back(make_int4(b, s, std::min(e1, s + chunk), nsplit > 1 ? 1 : 0));
    }
    while (items.size() % kXcds) items.push_back(make_int4(0, 0, 0, 0));
    return items;
}

// Target blocks of the matrix-core M2L: boxes of level L with the same octant, TB at a
// time in Morton order; per block the displacements present (sorted by rotation and
// distance) with the source box of every target (-1 if that target lacks it).
void build_m2l_blocks(hfmm_plan& P, int L, const std::vector<int>& targets, const std::vector<int>& ptr,
                      const std::vector<int>& src, const std::vector<unsigned short>& dir,
                      const std::vector<double>& coax_d, int n_dist, const std::vector<DispInfo>& dinfo) {
    if (P.m2l_kernel == 1) return;
    const int p = P.p_level[L];
    if (p > kM2LBlockMaxP) return;   // higher degrees use the generic kernel (coarse levels only)
    const int fp32 = P.m2l_fp32[L];
    const int eb = fp32 ? 4 : 8;
    // the instantiation's stage-3 units per wave (FP32 one-tile instances pack small degrees)
    const int maxu = fp32 && m2l_tiles(p + 1) == 1 ? m2l_block_maxu_packed(p) : m2l_block_maxu(p);
    // shape of one operator application: TB target boxes x G directions (of equal theta
    // and |d|).  A requested pair count TB*G (m2l_kernel = 8, 16, 32, 64) takes the first
    // fitting shape with that count in the order below (larger TB first: partly filled
    // direction groups waste columns, target blocks rarely do).  Otherwise, from
    // measurements at N = 1M, k = 40 (finest level, 5.4M pairs): FP32 takes 16 x 2 up to
    // p = 11 (5.27 ms at p = 11 against 5.61 for 8 x 2) and 8 x 2 above (see below); FP64
    // takes 8 x 1 (38.9 ms at p = 17 against 48.4 for 8 x 2); then the next smaller shape
    // that fits.
    auto fits = [&](int c, int g) {
        if (!fp32 && !(c == 8 || (c == 16 && g == 1))) return false;   // FP64 instantiations
        if (m2l_acc_regs(maxu, c, eb) > kM2LMaxAccRegs) return false;
        return m2l_block_lds_bytes(p, c, g, eb) <= P.max_lds;
    };
    int tb = 0, grp = 0;
    const int shapes[6][2] = {{16, 4}, {16, 2}, {8, 4}, {8, 2}, {16, 1}, {8, 1}};
    if (P.m2l_kernel > 1)
        for (auto& sh : shapes)
            if (!tb && sh[0] * sh[1] == P.m2l_kernel && fits(sh[0], sh[1])) tb = sh[0], grp = sh[1];
    // with the packed stage-3 units the one-tile FP32 instances fit 16 x 2 up to p = 15:
    // at p = 12 on the finest level (32768 targets) 5.20 ms against 6.0 for 8 x 2, but on
    // the 4096 targets of level 4 0.96 against 0.80 ms (fewer, longer work items)
    const int pref[3][2][2] = {{{16, 2}, {8, 2}}, {{8, 2}, {8, 1}}, {{8, 1}, {8, 2}}};
    const bool wide = p <= 11 || (p <= 15 && targets.size() >= 16384);
    const int* pr = &pref[!fp32 ? 2 : wide ? 0 : 1][0][0];
    for (int i = 0; i < 2 && !tb; ++i)
        if (fits(pr[2 * i], pr[2 * i + 1])) tb = pr[2 * i], grp = pr[2 * i + 1];
    for (auto& sh : shapes)
        if (!tb && fits(sh[0], sh[1])) tb = sh[0], grp = sh[1];
    if (!tb) return;
    std::vector<int> units1, units2, units3;
    make_units(p, p, units1, units2, units3, fp32 && m2l_tiles(p + 1) == 1);

    const Tree& t = P.tree;
    std::vector<int> tpos(t.n_boxes(), -1);     // target ordinal in `targets`
    for (size_t i = 0; i < targets.size(); ++i) tpos[targets[i]] = (int)i;
    std::vector<int> blk_targets, blk_ptr{0}, grp_off;
    std::vector<unsigned short> grp_dir;
    for (int o = 0; o < 8; ++o) {
        std::vector<int> boxes;
        for (int b = t.level_start[L]; b < t.level_start[L + 1]; ++b)
            if ((int)(t.code[b] & 7) == o && tpos[b] >= 0) boxes.push_back(b);
        for (size_t b0 = 0; b0 < boxes.size(); b0 += tb) {
            // (theta, dist) -> direction -> sources of the TB targets
            std::map<std::pair<int, int>, std::map<int, std::vector<int>>> ents;
            for (int j = 0; j < tb; ++j) {
                const int b = b0 + j < boxes.size() ? boxes[b0 + j] : -1;
                blk_targets.push_back(b);
                if (b < 0) continue;
                const int ti = tpos[b];
                for (int e = ptr[ti]; e < ptr[ti + 1]; ++e) {
                    const int d = dir[e];
                    auto& byd = ents[std::make_pair(dinfo[d].theta, dinfo[d].dist)];
                    auto it = byd.find(d);
                    if (it == byd.end()) it = byd.emplace(d, std::vector<int>(tb, -1)).first;
                    it->second[j] = (int)P.exp_off[src[e]];
                }
            }
            for (auto& kv : ents) {
                std::vector<std::pair<int, const std::vector<int>*>> ds;
                for (auto& dv : kv.second) ds.push_back({dv.first, &dv.second});
                for (size_t g0 = 0; g0 < ds.size(); g0 += grp)
                    for (int g = 0; g < grp; ++g) {
                        const bool has = g0 + g < ds.size();
                        grp_dir.push_back(has ? (unsigned short)ds[g0 + g].first : (unsigned short)0xFFFF);
                        for (int j = 0; j < tb; ++j) grp_off.push_back(has ? (*ds[g0 + g].second)[j] : -1);
                    }
            }
            blk_ptr.push_back((int)(grp_dir.size() / grp));
        }
    }
    const std::vector<int4> items = make_items(t, blk_targets, blk_ptr, tb, true);
    M2LBlockLevel& B = P.m2lb[L];
    B.use = true;
    B.fp32 = fp32;
    B.tb = tb;
    B.grp = grp;
    B.ps = B.pt = p;
    B.src_level = B.tgt_level = L;
    B.n_items = (int)items.size();
    B.items = P.upload(items);
    B.targets = P.upload(blk_targets);
    B.dir = P.upload(grp_dir);
    B.off = P.upload(grp_off);
    B.units1 = P.upload(units1);
    B.n_units1 = (int)units1.size();
    B.units2 = P.upload(units2);
    B.n_units2 = (int)units2.size();
    B.units3 = P.upload(units3);
    B.n_units3 = (int)units3.size();
    std::vector<double> cf;
    std::vector<int> coff, coffq;
    B.coax_stride = coax_fragments(coax_d, n_dist, p, p, cf, coff, coffq);
    if (fp32) {
        std::vector<float> c32(cf.begin(), cf.end());
        B.coax = P.upload(c32);
    } else {
        B.coax = P.upload(cf);
    }
    B.coax_off = P.upload(coff);
    B.coax_offq = P.upload(coffq);
    prepare_m2l_block(p, tb, grp, fp32);
}

// M2M (kind 1: the parents of level L from their children) and L2L (kind 2: the children
// at level L+1 from their parents) on the matrix-core block kernel, in FP32 or FP64 (see
// below; the results are FP64 either way).  Blocks of 8 targets,
// one direction per group: an M2M block runs one group per child octant present among
// its parents, every target receiving all of its children within the one work item, so
// the result is stored (no clear of the parent multipoles is needed); an L2L block holds
// 8 children of one octant and a single group (their parents' locals), added to what
// M2L left there.  tab: the coaxial table (pt+1-m) x (ps+1-m) per m of this level pair.
void build_tree_blocks(hfmm_plan& P, int L, int kind, const std::vector<double>& tab) {
    if (P.m2l_kernel == 1 || !P.d_rf1) return;
    const Tree& t = P.tree;
    const int ps = kind == 1 ? P.p_level[L + 1] : P.p_level[L];
    const int pt = kind == 1 ? P.p_level[L] : P.p_level[L + 1];
    const int pm = std::max(ps, pt);
    if (pm > kM2LBlockMaxP) return;
    // FP32 arithmetic (FP64 results) where the M2L of every level from 2 to L+1 runs in FP32:
    // the translated expansions then only feed FP32 translations, whose own rounding is of
    // the same size; a chain through an FP64 M2L level keeps FP64.  Split M2M levels (below)
    // keep FP64 as well: there the FP32 form was slower (N=1M, k=40: 55.7 against 40.4 us
    // for level 4 -> 3, where the level 5 -> 4 M2M took 81.9 against 113.5 us).
    bool fp32 = P.d_rf1_32 != nullptr;
    for (int l = 2; l <= L + 1; ++l) fp32 = fp32 && P.m2l_fp32[l] && P.p_level[l] <= kM2LBlockMaxP;
    const int tb = 8, grp = 1;
    if (m2l_block_lds_bytes(pm, tb, grp, 8) > P.max_lds) return;
    std::vector<int> blk_targets, blk_ptr{0}, grp_off;
    std::vector<unsigned short> grp_dir;
    if (kind == 1) {
        // parents with children, in level order, and their child per octant
        std::vector<int> parents;
        std::vector<std::array<int, 8>> child;
        for (int c = t.level_start[L + 1]; c < t.level_start[L + 2]; ++c) {
            const int par = t.parent[c];
            if (parents.empty() || parents.back() != par) {
                parents.push_back(par);
                child.push_back({-1, -1, -1, -1, -1, -1, -1, -1});
            }
            child.back()[octant(t.code[c])] = c;
        }
        for (size_t b0 = 0; b0 < parents.size(); b0 += tb) {
            for (int j = 0; j < tb; ++j) blk_targets.push_back(b0 + j < parents.size() ? parents[b0 + j] : -1);
            for (int o = 0; o < 8; ++o) {
                bool any = false;
                for (int j = 0; j < tb && b0 + j < parents.size(); ++j) any |= child[b0 + j][o] >= 0;
                if (!any) continue;
                grp_dir.push_back((unsigned short)(kDirM2M + o));
                for (int j = 0; j < tb; ++j) {
                    const int c = b0 + j < parents.size() ? child[b0 + j][o] : -1;
                    grp_off.push_back(c >= 0 ? (int)P.exp_off[c] : -1);
                }
            }
            blk_ptr.push_back((int)grp_dir.size());
        }
    } else {
        for (int o = 0; o < 8; ++o) {
            std::vector<int> kids;
            for (int c = t.level_start[L + 1]; c < t.level_start[L + 2]; ++c)
                if (octant(t.code[c]) == o) kids.push_back(c);
            for (size_t b0 = 0; b0 < kids.size(); b0 += tb) {
                grp_dir.push_back((unsigned short)(kDirL2L + o));
                for (int j = 0; j < tb; ++j) {
                    const int c = b0 + j < kids.size() ? kids[b0 + j] : -1;
                    blk_targets.push_back(c);
                    grp_off.push_back(c >= 0 ? (int)P.exp_off[t.parent[c]] : -1);
                }
                blk_ptr.push_back((int)grp_dir.size());
            }
        }
    }
    if (blk_ptr.size() < 2) return;
    // coarse M2M levels have few blocks: their groups (child octants) are split over items
    // that add atomically, after the parents are cleared; L2L blocks hold one group
    const int nblk = (int)blk_ptr.size() - 1;
    const bool split = kind == 1 && nblk < 512;
    if (split) fp32 = false;
    const std::vector<int4> items = make_items(t, blk_targets, blk_ptr, tb, split);
    std::vector<int> units1, units2, units3;
    make_units(ps, pt, units1, units2, units3, fp32 && m2l_tiles(pm + 1) == 1);
    M2LBlockLevel& B = kind == 1 ? P.m2mb[L] : P.l2lb[L];
    B.use = true;
    B.fp32 = fp32;
    B.tb = tb;
    B.grp = grp;
    B.ps = ps;
    B.pt = pt;
    B.src_level = kind == 1 ? L + 1 : L;
    B.tgt_level = kind == 1 ? L : L + 1;
    B.n_items = (int)items.size();
    B.items = P.upload(items);
    B.targets = P.upload(blk_targets);
    B.dir = P.upload(grp_dir);
    B.off = P.upload(grp_off);
    B.units1 = P.upload(units1);
    B.n_units1 = (int)units1.size();
    B.units2 = P.upload(units2);
    B.n_units2 = (int)units2.size();
    B.units3 = P.upload(units3);
    B.n_units3 = (int)units3.size();
    B.split = split;
    if (split) {
        std::vector<int> z;
        for (int b : blk_targets)
            if (b >= 0) z.push_back(b);
        B.n_zero = (int)z.size();
        B.zero = P.upload(z);
    }
    std::vector<double> cf;
    std::vector<int> coff, coffq;
    B.coax_stride = coax_fragments(tab, 1, ps, pt, cf, coff, coffq);
    if (fp32) {
        std::vector<float> c32(cf.begin(), cf.end());
        B.coax = P.upload(c32);
    } else {
        B.coax = P.upload(cf);
    }
    B.coax_off = P.upload(coff);
    B.coax_offq = P.upload(coffq);
    prepare_m2l_block(pm, tb, grp, fp32);
}

void build_tables(hfmm_plan& P) {
    const Tree& t = P.tree;
    // directions
    std::map<std::pair<int, int>, int> theta_idx;
    std::vector<double> thetas;
    std::vector<DispInfo> dinfo(kNumDirs, DispInfo{0, 0});
    std::vector<double> phis(kNumDirs, 0.0);
    std::map<int, int> dist_idx;       // |d|^2 -> index
    std::vector<int> dist_d2;
    auto theta_of = [&](int dz, int d2) {
        auto key = std::make_pair(dz, d2);
        auto it = theta_idx.find(key);
        if (it != theta_idx.end()) return it->second;
        int id = (int)thetas.size();
        thetas.push_back(std::acos(std::max(-1.0, std::min(1.0, dz / std::sqrt((double)d2)))));
        theta_idx[key] = id;
        return id;
    };
    for (int dx = -3; dx <= 3; ++dx)
        for (int dy = -3; dy <= 3; ++dy)
            for (int dz = -3; dz <= 3; ++dz) {
                const int d = (dx + 3) * 49 + (dy + 3) * 7 + (dz + 3);
                if (std::max(std::abs(dx), std::max(std::abs(dy), std::abs(dz))) < 2) continue;
                const int d2 = dx * dx + dy * dy + dz * dz;
                if (!dist_idx.count(d2)) {
                    dist_idx[d2] = (int)dist_d2.size();
                    dist_d2.push_back(d2);
                }
                dinfo[kDirV + d] = DispInfo{theta_of(dz, d2), dist_idx[d2]};
                phis[kDirV + d] = std::atan2((double)dy, (double)dx);
            }
    for (int o = 0; o < 8; ++o) {
        const int sx = 2 * (o & 1) - 1, sy = 2 * ((o >> 1) & 1) - 1, sz = 2 * ((o >> 2) & 1) - 1;
        dinfo[kDirM2M + o] = DispInfo{theta_of(-sz, 3), 0};
        phis[kDirM2M + o] = std::atan2((double)-sy, (double)-sx);
        dinfo[kDirL2L + o] = DispInfo{theta_of(sz, 3), 0};
        phis[kDirL2L + o] = std::atan2((double)sy, (double)sx);
    }
    const int pmax = P.pmax;
    const int rsz = rot_size(pmax);
    std::vector<double> rot((size_t)thetas.size() * rsz);
    for (size_t i = 0; i < thetas.size(); ++i) rotation_table(thetas[i], pmax, &rot[i * rsz]);
    P.d_rot = P.upload(rot);
    if (P.m2l_kernel != 1) {
        // rotation blocks n <= min(pmax, kM2LBlockMaxP) in the (u, v) basis as A fragments:
        // rf1 holds the blocks of E_n^T (forward rotation; u and v in their LDS order),
        // rf3 those of E_n (back rotation; output rows of both blocks indexed by the order
        // m = 0..n, the v block's row 0 zero, inputs in LDS order)
        const int pb = std::min(pmax, kM2LBlockMaxP);
        std::vector<int> off1(2 * (pb + 1), 0), off3(2 * (pb + 1), 0);
        int s1 = 0, s3 = 0;
        for (int n = 0; n <= pb; ++n) {
            off1[2 * n] = s1;
            s1 += m2l_tiles(n + 1) * m2l_tiles(n + 1) * 256;
            off1[2 * n + 1] = s1;
            s1 += m2l_tiles(n) * m2l_tiles(n) * 256;
            off3[2 * n] = s3;
            s3 += m2l_tiles(n + 1) * m2l_tiles(n + 1) * 256;
            off3[2 * n + 1] = s3;
            s3 += m2l_tiles(n + 1) * m2l_tiles(n) * 256;
        }
        // blocks of at most kM2LQuadMaxRows rows also get fragments for the 4x4x1 MFMA with its A broadcast
        // from lanes 0..3 (per 4-row tile t and k group g, t-major: lane l holds row
        // 4t + l % 4, k = 4g..4g+3), after the 16x16 fragments
        std::vector<int> off1q(2 * (pb + 1), -1);
        for (int n = 0; n <= pb; ++n)
            for (int h = 0; h < 2; ++h) {
                const int len = h ? n : n + 1;
                if (len >= 1 && len <= kM2LQuadMaxRows) {
                    off1q[2 * n + h] = s1;
                    s1 += ((len + 3) / 4) * ((len + 3) / 4) * 256;
                }
            }
        // packed stage-3 units (n1 < n2, n1 + n2 <= kM2LPackMax): a 16 x 16 block-diagonal
        // fragment each for u (rows 0..n1 and 8..8+n2; inputs n1's n1+1, then n2's n2+1) and
        // for v (same rows; inputs n1's n1, then n2's n2)
        std::vector<int> off3p(64, -1);
        for (int n1 = 0; n1 <= kM2LPackMax; ++n1)
            for (int n2 = n1 + 1; n1 + n2 <= kM2LPackMax; ++n2)
                if (n2 <= pb) {
                    off3p[8 * n1 + n2] = s3;
                    s3 += 512;
                }
        std::vector<double> f1((size_t)thetas.size() * s1, 0.0), f3((size_t)thetas.size() * s3, 0.0);
        // A fragments of an R x K matrix a(i, k) at dst
        auto put = [](double* dst, int R, int K, const std::function<double(int, int)>& a) {
            const int nk = m2l_tiles(K);
            for (int mt = 0; mt < m2l_tiles(R); ++mt)
                for (int kg = 0; kg < nk; ++kg)
                    for (int lane = 0; lane < 64; ++lane)
                        for (int j = 0; j < 4; ++j) {
                            const int i = 16 * mt + (lane & 15), k = 16 * kg + 4 * j + (lane >> 4);
                            if (i < R && k < K) dst[((size_t)(mt * nk + kg) * 64 + lane) * 4 + j] = a(i, k);
                        }
        };
        // 4x4x1 fragments of an R x K matrix (R, K <= 8), 4-row tiles
        auto put4 = [](double* dst, int R, int K, const std::function<double(int, int)>& a) {
            const int nt4 = (R + 3) / 4, nk4 = (K + 3) / 4;
            for (int t4 = 0; t4 < nt4; ++t4)
                for (int g4 = 0; g4 < nk4; ++g4)
                    for (int lane = 0; lane < 64; ++lane)
                        for (int j = 0; j < 4; ++j) {
                            const int i = 4 * t4 + (lane & 3), k = 4 * g4 + j;
                            if (i < R && k < K) dst[((t4 * nk4 + g4) * 64 + lane) * 4 + j] = a(i, k);
                        }
        };
        std::vector<double> Wu, Wv;
        std::vector<std::vector<double>> Wu3(pb + 1), Wv3(pb + 1);   // stage-3 blocks of one rotation
        for (size_t th = 0; th < thetas.size(); ++th) {
            for (int n = 0; n <= pb; ++n) {
                const double* E = &rot[th * rsz + rot_offset(n)];
                for (int st = 0; st < 2; ++st) {
                    rotation_uv_blocks(E, n, st == 0, Wu, Wv);   // stage 1: E^T, stage 3: E
                    auto wu = [&](int i, int k) { return Wu[(size_t)i * (n + 1) + k]; };
                    // v index m = 1..n; LDS order k' = n - m
                    auto wv = [&](int mi, int mk) { return Wv[(size_t)(mi - 1) * n + (mk - 1)]; };
                    if (st == 0) {
                        double* b = &f1[th * s1];
                        put(b + off1[2 * n], n + 1, n + 1, wu);
                        if (n > 0) put(b + off1[2 * n + 1], n, n, [&](int i, int k) { return wv(n - i, n - k); });
                        if (off1q[2 * n] >= 0) put4(b + off1q[2 * n], n + 1, n + 1, wu);
                        if (off1q[2 * n + 1] >= 0)
                            put4(b + off1q[2 * n + 1], n, n, [&](int i, int k) { return wv(n - i, n - k); });
                    } else {
                        double* b = &f3[th * s3];
                        put(b + off3[2 * n], n + 1, n + 1, wu);
                        if (n > 0)
                            put(b + off3[2 * n + 1], n + 1, n,
                                [&](int i, int k) { return i == 0 ? 0.0 : wv(i, n - k); });
                        Wu3[n] = Wu;
                        Wv3[n] = Wv;
                    }
                }
            }
            double* b3 = &f3[th * s3];
            for (int n1 = 0; n1 <= kM2LPackMax; ++n1)
                for (int n2 = n1 + 1; n1 + n2 <= kM2LPackMax; ++n2) {
                    if (off3p[8 * n1 + n2] < 0) continue;
                    const std::vector<double>&U1 = Wu3[n1], &U2 = Wu3[n2], &V1 = Wv3[n1], &V2 = Wv3[n2];
                    put(b3 + off3p[8 * n1 + n2], 16, n1 + n2 + 2, [&](int i, int k) {
                        if (i < 8) return i <= n1 && k <= n1 ? U1[(size_t)i * (n1 + 1) + k] : 0.0;
                        const int ii = i - 8, kk = k - n1 - 1;
                        return ii <= n2 && kk >= 0 ? U2[(size_t)ii * (n2 + 1) + kk] : 0.0;
                    });
                    // v: row m (0: zero), input k in LDS order (v_m at k = n - m)
                    put(b3 + off3p[8 * n1 + n2] + 256, 16, n1 + n2, [&](int i, int k) {
                        if (i < 8) {
                            if (i == 0 || i > n1 || k >= n1) return 0.0;
                            return V1[(size_t)(i - 1) * n1 + (n1 - k - 1)];
                        }
                        const int ii = i - 8, kk = k - n1;
                        if (ii == 0 || ii > n2 || kk < 0) return 0.0;
                        return V2[(size_t)(ii - 1) * n2 + (n2 - kk - 1)];
                    });
                }
        }
        P.rf_stride1 = s1;
        P.rf_stride3 = s3;
        P.d_rf_off1 = P.upload(off1);
        P.d_rf_off1q = P.upload(off1q);
        P.d_rf_off3 = P.upload(off3);
        P.d_rf_off3p = P.upload(off3p);
        if (P.precision == HFMM_PREC_MIXED) {
            P.d_rf1_32 = P.upload(std::vector<float>(f1.begin(), f1.end()));
            P.d_rf3_32 = P.upload(std::vector<float>(f3.begin(), f3.end()));
        }
        P.d_rf1 = P.upload(f1);   // FP64: M2L, M2M and L2L of FP64 levels
        P.d_rf3 = P.upload(f3);
    }
    std::vector<double2> phase((size_t)kNumDirs * (pmax + 1));
    for (int d = 0; d < kNumDirs; ++d)
        for (int m = 0; m <= pmax; ++m)
            phase[(size_t)d * (pmax + 1) + m] = make_double2(std::cos(m * phis[d]), std::sin(m * phis[d]));
    P.d_phase = P.upload(phase);
    if (P.precision == HFMM_PREC_MIXED) {
        std::vector<float2> ph32(phase.size());
        for (size_t i = 0; i < phase.size(); ++i) ph32[i] = make_float2((float)phase[i].x, (float)phase[i].y);
        P.d_phase32 = P.upload(ph32);
    }
    P.d_dinfo = P.upload(dinfo);

    const Gaunt G = make_gaunt(pmax);
    const int D = P.depth;
    // V lists grouped by level / target
    {
        const auto& vt = P.lists.v_tgt;
        const auto& vs = P.lists.v_src;
        for (int L = 2; L <= D; ++L) {
            std::vector<int> targets, ptr{0}, src;
            std::vector<unsigned short> dir;
            for (size_t e = 0; e < vt.size(); ++e) {
                const int tb = vt[e];
                if (t.level[tb] != L) continue;
                if (targets.empty() || targets.back() != tb) {
                    if (!targets.empty()) ptr.push_back((int)src.size());
                    targets.push_back(tb);
                }
                const int sb = vs[e];
                int tx, ty, tz, sx, sy, sz;
                morton_grid(t.code[tb], tx, ty, tz);
                morton_grid(t.code[sb], sx, sy, sz);
                const int dx = tx - sx, dy = ty - sy, dz = tz - sz;
                src.push_back(sb);
                dir.push_back((unsigned short)(kDirV + (dx + 3) * 49 + (dy + 3) * 7 + (dz + 3)));
            }
            if (targets.empty()) continue;
            ptr.push_back((int)src.size());
            XlateLaunch& X = P.m2l[L];
            X.n_targets = (int)targets.size();
            X.targets = P.upload(targets);
            X.ptr = P.upload(ptr);
            X.src = P.upload(src);
            X.dir = P.upload(dir);
            X.ps = X.pt = P.p_level[L];
            X.src_level = X.tgt_level = L;
            X.coax_stride = coax_size(X.pt, X.ps);
            std::vector<double> cx((size_t)dist_d2.size() * X.coax_stride * 2);
            for (size_t i = 0; i < dist_d2.size(); ++i)
                coax_table(0, G, P.k, 2.0 * P.h[L] * std::sqrt((double)dist_d2[i]), P.R[L], P.R[L], X.ps, X.pt,
                           &cx[i * X.coax_stride * 2]);
            X.coax = (double2*)P.upload(cx);
            build_m2l_blocks(P, L, targets, ptr, src, dir, cx, (int)dist_d2.size(), dinfo);
        }
    }
    // M2M (parent level L <- children at L+1) and L2L (parent level L -> children at L+1)
    for (int L = 2; L < D; ++L) {
        std::vector<int> targets, ptr{0}, src;
        std::vector<unsigned short> dir;
        std::vector<int> ctargets, cptr{0}, csrc;
        std::vector<unsigned short> cdir;
        for (int c = t.level_start[L + 1]; c < t.level_start[L + 2]; ++c) {
            const int par = t.parent[c];
            if (targets.empty() || targets.back() != par) {
                if (!targets.empty()) ptr.push_back((int)src.size());
                targets.push_back(par);
            }
            src.push_back(c);
            dir.push_back((unsigned short)(kDirM2M + octant(t.code[c])));
            ctargets.push_back(c);
            csrc.push_back(par);
            cdir.push_back((unsigned short)(kDirL2L + octant(t.code[c])));
            cptr.push_back((int)csrc.size());
        }
        if (targets.empty()) continue;
        ptr.push_back((int)src.size());
        XlateLaunch& A = P.m2m[L];
        A.n_targets = (int)targets.size();
        A.targets = P.upload(targets);
        A.ptr = P.upload(ptr);
        A.src = P.upload(src);
        A.dir = P.upload(dir);
        A.ps = P.p_level[L + 1];
        A.pt = P.p_level[L];
        A.src_level = L + 1;
        A.tgt_level = L;
        A.coax_stride = coax_size(A.pt, A.ps);
        std::vector<double> cx((size_t)A.coax_stride * 2);
        coax_table(1, G, P.k, P.R[L + 1], P.R[L + 1], P.R[L], A.ps, A.pt, cx.data());
        A.coax = (double2*)P.upload(cx);
        build_tree_blocks(P, L, 1, cx);
        XlateLaunch& B = P.l2l[L];
        B.n_targets = (int)ctargets.size();
        B.targets = P.upload(ctargets);
        B.ptr = P.upload(cptr);
        B.src = P.upload(csrc);
        B.dir = P.upload(cdir);
        B.ps = P.p_level[L];
        B.pt = P.p_level[L + 1];
        B.src_level = L;
        B.tgt_level = L + 1;
        B.coax_stride = coax_size(B.pt, B.ps);
        std::vector<double> cy((size_t)B.coax_stride * 2);
        coax_table(2, G, P.k, P.R[L + 1], P.R[L], P.R[L + 1], B.ps, B.pt, cy.data());
        B.coax = (double2*)P.upload(cy);
        build_tree_blocks(P, L, 2, cy);
    }
}

P2PArgs p2p_args(const hfmm_plan& P, int out_mode, double2* out);

void build_p2p(hfmm_plan& P) {
    const Tree& t = P.tree;
    const Lists& ls = P.lists;
    const int nb = t.n_boxes();
    auto csr = [nb](const std::vector<int>& tg, const std::vector<int>& sr, std::vector<int>& ptr, std::vector<int>& s) {
        ptr.assign(nb + 1, 0);
        for (int x : tg) ptr[x + 1]++;
        for (int i = 0; i < nb; ++i) ptr[i + 1] += ptr[i];
        s.assign(sr.size(), 0);
        std::vector<int> pos(ptr.begin(), ptr.end() - 1);
        for (size_t e = 0; e < tg.size(); ++e) s[pos[tg[e]]++] = sr[e];
    };
    std::vector<int> up, us, wp, ws, xp, xs;
    csr(ls.u_tgt, ls.u_src, up, us);
    csr(ls.w_tgt, ls.w_src, wp, ws);
    csr(ls.x_tgt, ls.x_src, xp, xs);
    // box centre and half diagonal, for the bound on k r over a leaf's pairs
    auto geom = [&](int b, double c[3]) {
        int g[3];
        morton_grid(t.code[b], g[0], g[1], g[2]);
        const double w = t.extent * std::ldexp(1.0, -t.level[b]);
        for (int a = 0; a < 3; ++a) c[a] = (g[a] + 0.5) * w + t.pmin[a];
        return 0.5 * std::sqrt(3.0) * w;
    };
    // leaves whose pairs stay within k r <= kP2PFp32MaxKr first (FP32 near field), then the
    // rest (FP32 with the phase in FP64); Morton order within each class
    std::vector<int> leaves[2];
    double kr_max = 0.0;
    for (int b = 0; b < nb; ++b) {
        if (!t.is_leaf[b]) continue;
        double kr = 0.0;
        if (P.precision == HFMM_PREC_MIXED) {
            double cb[3], cs[3];
            const double hb = geom(b, cb);
            auto reach = [&](int box) {
                const double hs = geom(box, cs);
                const double d = std::sqrt((cb[0] - cs[0]) * (cb[0] - cs[0]) + (cb[1] - cs[1]) * (cb[1] - cs[1]) +
                                           (cb[2] - cs[2]) * (cb[2] - cs[2]));
                kr = std::max(kr, P.k * (d + hb + hs));
            };
            reach(b);
            for (int e = up[b]; e < up[b + 1]; ++e) reach(us[e]);
            for (int e = wp[b]; e < wp[b + 1]; ++e) reach(ws[e]);
            for (int a = b; a >= 0; a = t.parent[a])
                for (int e = xp[a]; e < xp[a + 1]; ++e) reach(xs[e]);
        }
        leaves[kr > kP2PFp32MaxKr].push_back(b);
        kr_max = std::max(kr_max, kr);
    }
    P.n_p2p_fp32 = (int)leaves[0].size();
    P.p2p_lowk = P.precision == HFMM_PREC_MIXED && kr_max <= kP2PLowKr;
    std::vector<int> leaf_box, rng_ptr{0}, rng_start, rng_pref, total;
    std::vector<double> floor2;
    std::vector<int> far_leaves;
    int64_t pairs = 0;
    for (int b = 0; b < nb; ++b)
        if (t.is_leaf[b] && t.level[b] >= 2) far_leaves.push_back(b);
    for (int cls = 0; cls < 2; ++cls)
        for (int b : leaves[cls]) {
            leaf_box.push_back(b);
            int acc = 0;
            double src_maxabs = 0.0;
            auto add = [&](int box) {
                rng_start.push_back(t.pstart[box]);
                rng_pref.push_back(acc);
                acc += t.count[box];
                src_maxabs = std::max(src_maxabs, P.box_maxabs[box]);
            };
            add(b);
            for (int e = up[b]; e < up[b + 1]; ++e) add(us[e]);
            for (int e = wp[b]; e < wp[b + 1]; ++e) add(ws[e]);
            for (int a = b; a >= 0; a = t.parent[a])
                for (int e = xp[a]; e < xp[a + 1]; ++e) add(xs[e]);
            rng_ptr.push_back((int)rng_start.size());
            total.push_back(acc);
            // the reference's coincidence floor (jaxfmm/near.py:252-261, "pair" form in FP64):
            // pairs with r <= 16 eps (max |x_t| + max |x_s|) contribute nothing
            const double fl = 16.0 * std::numeric_limits<double>::epsilon() * (P.box_maxabs[b] + src_maxabs);
            floor2.push_back(fl * fl);
            pairs += (int64_t)acc * t.count[b];
        }
    P.n_leaves = (int)leaf_box.size();
    P.n_far_leaves = (int)far_leaves.size();
    P.n_p2p_pairs = pairs;
    P.d_leaf_box = P.upload(leaf_box);
    P.d_far_leaves = P.upload(far_leaves);
    P.d_rng_ptr = P.upload(rng_ptr);
    P.d_rng_start = P.upload(rng_start);
    P.d_rng_pref = P.upload(rng_pref);
    P.d_leaf_src_total = P.upload(total);
    P.d_floor2 = P.upload(floor2);
    P.box_maxabs = std::vector<double>();
    // FP32-phase leaves whose estimated error from FP32 coordinate differences is within
    // p2p_coord32_tol take them (the check runs once, on the GPU, in FP64)
    if (P.precision == HFMM_PREC_MIXED && !P.p2p_lowk && P.p2p_coord32_tol > 0.0 && P.n_p2p_fp32 > 0) {
        P.d_coord32 = P.alloc<unsigned char>(P.n_leaves);
        HIPCHK(hipMemset(P.d_coord32, 0, P.n_leaves));
        launch_p2p_coord_check(p2p_args(P, 0, nullptr), P.n_p2p_fp32, P.p2p_coord32_tol, P.d_coord32, P.stream);
        std::vector<unsigned char> f(P.n_leaves);
        HIPCHK(hipMemcpyAsync(f.data(), P.d_coord32, P.n_leaves, hipMemcpyDeviceToHost, P.stream));
        HIPCHK(hipStreamSynchronize(P.stream));
        for (unsigned char v : f) P.n_leaves_coord32 += v;
    }
}

void build_plan(hfmm_plan& P, const double* xyz, int64_t n, double k, const hfmm_opts& o) {
    if (!std::isfinite(k)) fail(HFMM_ERR_ARG, "k must be a finite real number");
    if (n <= 0 || n > (int64_t)1 << 30) fail(HFMM_ERR_ARG, "positions must be a nonempty finite (N,3) array");
    for (int64_t i = 0; i < 3 * n; ++i)
        if (!std::isfinite(xyz[i])) fail(HFMM_ERR_ARG, "positions must be a nonempty finite (N,3) array");
    if (o.p < 2 || o.n_max < 1) fail(HFMM_ERR_ARG, "p must be an integer >=2; N_max must be a positive integer");
    if (o.p > kMaxP) fail(HFMM_ERR_ARG, "p exceeds the supported maximum of 40");
    for (int L = 0; L < HFMM_MAX_LEVELS; ++L)
        if (o.p_level[L] != 0 && (o.p_level[L] < 2 || o.p_level[L] > kMaxP))
            fail(HFMM_ERR_ARG, "p_level entries must be 0 or in [2, 40]");
    if (o.precision != HFMM_PREC_FP64 && o.precision != HFMM_PREC_MIXED) fail(HFMM_ERR_ARG, "unknown precision");
    if (o.m2l_kernel != 0 && o.m2l_kernel != 1 && o.m2l_kernel != 8 && o.m2l_kernel != 16 && o.m2l_kernel != 32 &&
        o.m2l_kernel != 64)
        fail(HFMM_ERR_ARG, "m2l_kernel must be 0, 1, 8, 16, 32 or 64");
    if (o.p2p_tile != 0 && o.p2p_tile != 1 && o.p2p_tile != 2 && o.p2p_tile != 4)
        fail(HFMM_ERR_ARG, "p2p_tile must be 0, 1, 2 or 4");
    check_device();
    upload_legendre_coefficients();
    P.n = n;
    P.k = std::fabs(k);
    P.conj = k < 0;
    P.precision = o.precision;
    P.use_graph = o.use_graph;
    P.m2l_kernel = o.m2l_kernel;
    P.p2p_tile = o.p2p_tile == 0 ? 2 : o.p2p_tile;
    if (!std::isfinite(o.p2p_coord32_tol)) fail(HFMM_ERR_ARG, "p2p_coord32_tol must be a finite number");
    P.p2p_coord32_tol = o.p2p_coord32_tol == 0.0 ? HFMM_P2P_COORD32_TOL : o.p2p_coord32_tol;
    {
        int dev = 0, lds = 65536;
        HIPCHK(hipGetDevice(&dev));
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
            P.max_lds = (size_t)lds;
    }
    if (o.tree_device) {
        build_tree_device(xyz, n, o.n_max, o.max_depth, P.tree, P.lists);
    } else {
        P.tree = build_tree(xyz, n, o.n_max, o.max_depth);
        P.lists = build_lists(P.tree);
    }
    const Tree& t = P.tree;
    P.depth = t.depth;
    P.pmax = 0;
    for (int L = 0; L < kMaxLevels; ++L) {
        P.p_level[L] = (L < HFMM_MAX_LEVELS && o.p_level[L] > 0) ? o.p_level[L] : o.p;
        P.h[L] = t.extent * std::ldexp(1.0, -(L + 1));
        P.R[L] = std::sqrt(3.0) * P.h[L];
        if (L >= 1 && L <= t.depth) P.pmax = std::max(P.pmax, P.p_level[L]);
        P.p_eff[L] = P.p_level[L];
        const double kd = P.k * 2.0 * P.R[L];   // k times the box diagonal
        P.m2l_fp32[L] = o.precision == HFMM_PREC_MIXED &&
                        (o.m2l_fp64_levels < 0 ? kd <= HFMM_MIXED_FP32_MAX_KD : !((o.m2l_fp64_levels >> L) & 1));
    }
    HIPCHK(hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
    // sorted positions (SoA, FP64) and permutation
    std::vector<double> X(n), Y(n), Z(n);
    std::vector<int> perm(n);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t o2 = t.perm[i];
        X[i] = xyz[3 * o2];
        Y[i] = xyz[3 * o2 + 1];
        Z[i] = xyz[3 * o2 + 2];
        perm[i] = (int)o2;
    }
    P.dX = P.upload(X);
    P.dY = P.upload(Y);
    P.dZ = P.upload(Z);
    P.d_perm = P.upload(perm);
    // boxes
    const int nb = t.n_boxes();
    // largest |x| per box (points of leaves, then children into parents: boxes are in level
    // order, so children come after their parents), for the near field's coincidence floor
    P.box_maxabs.assign(nb, 0.0);
    for (int b = nb - 1; b >= 0; --b) {
        if (t.is_leaf[b])
            for (int i = t.pstart[b]; i < t.pstart[b] + t.count[b]; ++i)
                P.box_maxabs[b] = std::max(P.box_maxabs[b], std::sqrt(X[i] * X[i] + Y[i] * Y[i] + Z[i] * Z[i]));
        if (t.parent[b] >= 0) P.box_maxabs[t.parent[b]] = std::max(P.box_maxabs[t.parent[b]], P.box_maxabs[b]);
    }
    std::vector<double> centre(3 * (size_t)nb);
    P.exp_off.assign(nb, 0);
    int64_t off = 0;
    for (int b = 0; b < nb; ++b) {
        int gx, gy, gz;
        morton_grid(t.code[b], gx, gy, gz);
        const double s = t.extent * std::ldexp(1.0, -t.level[b]);
        const double half = t.extent * std::ldexp(1.0, -(t.level[b] + 1));
        centre[3 * b] = gx * s + half + t.pmin[0];
        centre[3 * b + 1] = gy * s + half + t.pmin[1];
        centre[3 * b + 2] = gz * s + half + t.pmin[2];
        P.exp_off[b] = off;
        off += ncoef(P.p_level[t.level[b]]);
    }
    P.n_coef = off;
    P.d_centre = P.upload(centre);
    P.d_level = P.upload(t.level);
    P.d_pstart = P.upload(t.pstart);
    P.d_count = P.upload(t.count);
    P.d_expoff = P.upload(P.exp_off);
    std::vector<int> pl(P.p_level, P.p_level + kMaxLevels);
    std::vector<double> rl(P.R, P.R + kMaxLevels);
    P.d_plevel = P.upload(pl);
    P.d_R = P.upload(rl);
    build_p2p(P);
    {
        // P2M / L2P descriptors of the far leaves, in the order of d_far_leaves
        std::vector<LeafDesc> ld;
        for (int b = 0; b < nb; ++b)
            if (t.is_leaf[b] && t.level[b] >= 2) {
                LeafDesc d{};
                d.box = b;
                d.level = t.level[b];
                d.pstart = t.pstart[b];
                d.count = t.count[b];
                d.exp_off = P.exp_off[b];
                d.cx = centre[3 * b];
                d.cy = centre[3 * b + 1];
                d.cz = centre[3 * b + 2];
                ld.push_back(d);
            }
        if ((int)ld.size() != P.n_far_leaves) fail(HFMM_ERR_INTERNAL, "far leaf descriptors out of step");
        P.d_leafdesc = P.upload(ld);
    }
    build_tables(P);
    // P2M: points staged per pass, sized for the largest leaf degree
    P.p_leafmax = 2;
    for (int b = 0; b < t.n_boxes(); ++b)
        if (t.is_leaf[b]) P.p_leafmax = std::max(P.p_leafmax, P.p_level[t.level[b]]);
    P.p2m_chunk = p2m_chunk(P.p_leafmax);
    P.dM = P.alloc<double2>(P.n_coef);
    bool any_fp32 = false, any_l2l32 = false;
    for (int L = 0; L < kMaxLevels; ++L) {
        any_fp32 = any_fp32 || (P.m2lb[L].use && P.m2lb[L].fp32) || (P.m2mb[L].use && P.m2mb[L].fp32);
        any_l2l32 = any_l2l32 || (P.l2lb[L].use && P.l2lb[L].fp32);
    }
    if (any_fp32) P.dM32 = P.alloc<float2>(P.n_coef);
    P.dL = P.alloc<double2>(P.n_coef);
    if (any_l2l32) P.dL32 = P.alloc<float2>(P.n_coef);
    P.d_q = P.alloc<double2>(n);
    P.d_far = P.alloc<double2>(n);
    P.d_qin = P.alloc<double2>(n);
    P.d_out = P.alloc<double2>(n);
    HIPCHK(hipDeviceSynchronize());
}

XlateArgs xargs(const hfmm_plan& P, const XlateLaunch& X, const double2* srcbuf, double2* dstbuf, int accumulate) {
    XlateArgs a;
    a.targets = X.targets;
    a.ptr = X.ptr;
    a.src = X.src;
    a.dir = X.dir;
    a.dinfo = P.d_dinfo;
    a.phase = P.d_phase;
    a.rot = P.d_rot;
    a.coax = X.coax;
    a.coax_stride = X.coax_stride;
    a.ps = P.p_eff[X.src_level];
    a.pt = P.p_eff[X.tgt_level];
    a.tps = X.ps;
    a.tpt = X.pt;
    a.pmax = P.pmax;
    a.accumulate = accumulate;
    a.split = 1;
    a.srcbuf = srcbuf;
    a.dstbuf = dstbuf;
    a.exp_off = P.d_expoff;
    return a;
}

M2LBlockArgs block_args(const hfmm_plan& P, const M2LBlockLevel& B, bool fp32, const void* src, double2* dst,
                        int overwrite) {
    M2LBlockArgs a;
    a.p = P.p_eff[B.src_level];
    a.p_tab = B.ps;
    a.pt = P.p_eff[B.tgt_level];
    a.pt_tab = B.pt;
    a.overwrite = overwrite;
    a.n_items = B.n_items;
    a.items = B.items;
    a.blk_targets = B.targets;
    a.grp_dir = B.dir;
    a.grp_off = B.off;
    a.units1 = B.units1;
    a.n_units1 = B.n_units1;
    a.units2 = B.units2;
    a.n_units2 = B.n_units2;
    a.units3 = B.units3;
    a.n_units3 = B.n_units3;
    a.rot1 = fp32 ? (const void*)P.d_rf1_32 : (const void*)P.d_rf1;
    a.rot3 = fp32 ? (const void*)P.d_rf3_32 : (const void*)P.d_rf3;
    a.rot_off1 = P.d_rf_off1;
    a.rot_off1q = P.d_rf_off1q;
    a.rot_off3 = P.d_rf_off3;
    a.rot_off3p = P.d_rf_off3p;
    a.rot_stride1 = P.rf_stride1;
    a.rot_stride3 = P.rf_stride3;
    a.phase = fp32 ? (const void*)P.d_phase32 : (const void*)P.d_phase;
    a.phase_stride = P.pmax + 1;
    a.dinfo = P.d_dinfo;
    a.coax = B.coax;
    a.coax_off = B.coax_off;
    a.coax_offq = B.coax_offq;
    a.coax_stride = B.coax_stride;
    a.M = src;
    a.L = dst;
    a.exp_off = P.d_expoff;
    return a;
}

// parts of one evaluate's device sequence (a batch runs them for all problems, with the
// block M2L and the near field as persistent launches over every problem)
constexpr int kSeqUp = 1;          // charges, clears, P2M, M2M
constexpr int kSeqM2LGeneric = 2;  // M2L levels on the generic kernel
constexpr int kSeqM2LBlock = 4;    // M2L levels on the matrix-core block kernel
constexpr int kSeqDown = 8;        // L2L, L2P
constexpr int kSeqP2P = 16;        // near field and output
constexpr int kSeqAll = 31;

// first coefficient of the expansions of tree level l (boxes are in level order); n_coef
// past the deepest level
int64_t level_coef(const hfmm_plan& P, int l) {
    const Tree& t = P.tree;
    if (l > t.depth) return P.n_coef;
    const int b = t.level_start[l];
    return b < t.n_boxes() ? P.exp_off[b] : P.n_coef;
}

int out_mode_of(const hfmm_plan& P, int q_type) {
    return (P.conj ? 1 : 0) | ((P.k == 0.0 && real_charges(q_type)) ? 2 : 0);
}

P2PArgs p2p_args(const hfmm_plan& P, int out_mode, double2* out) {
    P2PArgs pa{};
    pa.leaf_box = P.d_leaf_box;
    pa.bx = P.boxes();
    pa.rng_ptr = P.d_rng_ptr;
    pa.rng_start = P.d_rng_start;
    pa.rng_pref = P.d_rng_pref;
    pa.leaf_src_total = P.d_leaf_src_total;
    pa.floor2 = P.d_floor2;
    pa.X = P.dX;
    pa.Y = P.dY;
    pa.Z = P.dZ;
    pa.q = P.d_q;
    pa.k = P.k;
    pa.far = P.d_far;
    pa.perm = P.d_perm;
    pa.out_mode = out_mode;
    pa.out = out;
    pa.coord32 = P.d_coord32;
    return pa;
}

// The device sequence of one evaluate (the parts selected).  ev (optional, 8 entries)
// brackets the phases; lev (optional, 2 per level) brackets the M2L launch of every level.
// field (optional): E = -grad phi, 3 complex values per point in input order.
void run_sequence(hfmm_plan& P, const void* q, int q_type, double2* out, hipStream_t s, hipEvent_t* ev,
                  hipEvent_t* lev = nullptr, double2* field = nullptr, int parts = kSeqAll) {
    auto mark = [&](int i) {
        if (ev) HIPCHK(hipEventRecord(ev[i], s));
    };
    const DevBoxes bx = P.boxes();
    const DevLevels lv = P.levels();
    if (parts & kSeqUp) {
        mark(0);
        launch_gather_charges(q, q_type, P.d_perm, P.n, P.conj, P.d_q, s);
        launch_fill_zero(P.dL, P.n_coef, s);
        launch_fill_zero(P.d_far, P.n, s);
        mark(1);
        launch_p2m(P.d_leafdesc, P.n_far_leaves, P.p2m_chunk, lv, P.dX, P.dY, P.dZ, P.d_q, P.k, P.p_leafmax, P.dM, s);
        mark(2);
        bool conv[kMaxLevels + 1] = {};   // levels whose multipoles are in FP32 already
        for (int L = P.depth - 1; L >= 2; --L) {
            if (P.m2mb[L].use) {
                const M2LBlockLevel& B = P.m2mb[L];
                if (B.fp32) {   // the children's multipoles (complete now) in FP32
                    const int64_t c0 = level_coef(P, L + 1), c1 = level_coef(P, L + 2);
                    launch_to_float(P.dM + c0, P.dM32 + c0, c1 - c0, s);
                    conv[L + 1] = true;
                }
                if (B.split) launch_zero_boxes(B.zero, B.n_zero, P.d_expoff, ncoef(P.p_eff[L]), P.dM, s);
                launch_m2l_block(block_args(P, B, B.fp32, B.fp32 ? (const void*)P.dM32 : (const void*)P.dM, P.dM,
                                            B.split ? 0 : 1),
                                 B.tb, B.grp, B.fp32, s);
            }
            else if (P.m2m[L].n_targets)
                launch_xlate(xargs(P, P.m2m[L], P.dM, P.dM, 0), P.m2m[L].n_targets, s);
        }
        mark(3);
        // multipoles in FP32 for the FP32 block levels: every level not converted for an FP32
        // M2M (the converted ones need not be contiguous), one launch per run of levels
        bool any_fp32 = false;
        for (int L = 2; L <= P.depth; ++L) any_fp32 |= P.m2l[L].n_targets && P.m2lb[L].use && P.m2lb[L].fp32;
        if (any_fp32)
            for (int l0 = 0; l0 <= P.depth;) {
                if (conv[l0]) {
                    ++l0;
                    continue;
                }
                int l1 = l0;
                while (l1 + 1 <= P.depth && !conv[l1 + 1]) ++l1;
                const int64_t c0 = level_coef(P, l0), c1 = level_coef(P, l1 + 1);
                launch_to_float(P.dM + c0, P.dM32 + c0, c1 - c0, s);
                l0 = l1 + 1;
            }
    }
    for (int L = 2; L <= P.depth; ++L) {
        if (!P.m2l[L].n_targets) continue;
        const M2LBlockLevel& B = P.m2lb[L];
        if (!B.use) {
            if (!(parts & kSeqM2LGeneric)) continue;
            if (lev) HIPCHK(hipEventRecord(lev[2 * L], s));
            // coarse levels with few targets (degrees above the matrix-core kernel's): the
            // source lists split over up to 16 workgroups per target, so that about 2048
            // workgroups fill the GPU (level 2 of a uniform tree has 64 targets)
            XlateArgs a = xargs(P, P.m2l[L], P.dM, P.dL, 1);
            a.split = std::max(1, std::min(16, 2048 / std::max(1, P.m2l[L].n_targets)));
            launch_xlate(a, P.m2l[L].n_targets, s);
            if (lev) HIPCHK(hipEventRecord(lev[2 * L + 1], s));
            continue;
        }
        if (!(parts & kSeqM2LBlock)) continue;
        const M2LBlockArgs a =
            block_args(P, B, B.fp32, B.fp32 ? (const void*)P.dM32 : (const void*)P.dM, P.dL, 0);
        if (lev) HIPCHK(hipEventRecord(lev[2 * L], s));
        launch_m2l_block(a, B.tb, B.grp, B.fp32, s);
        if (lev) HIPCHK(hipEventRecord(lev[2 * L + 1], s));
    }
    if (parts & (kSeqM2LGeneric | kSeqM2LBlock)) mark(4);
    const int out_mode = out_mode_of(P, q_type);
    if (parts & kSeqDown) {
        for (int L = 2; L < P.depth; ++L) {
            const M2LBlockLevel& B = P.l2lb[L];
            if (B.use && B.fp32) {   // the parents' locals (complete now) in FP32
                const int64_t c0 = level_coef(P, L), c1 = level_coef(P, L + 1);
                launch_to_float(P.dL + c0, P.dL32 + c0, c1 - c0, s);
                launch_m2l_block(block_args(P, B, true, P.dL32, P.dL, 0), B.tb, B.grp, 1, s);
            } else if (B.use)
                launch_m2l_block(block_args(P, B, false, P.dL, P.dL, 0), B.tb, B.grp, 0, s);
            else if (P.l2l[L].n_targets)
                launch_xlate(xargs(P, P.l2l[L], P.dL, P.dL, 1), P.l2l[L].n_targets, s);
        }
        mark(5);
        if (field) {
            launch_fill_zero(P.d_far_e, 3 * P.n, s);
            launch_l2p_field(P.d_far_leaves, P.n_far_leaves, bx, lv, P.dX, P.dY, P.dZ, P.dL, P.k, P.p_leafmax,
                             P.d_far, P.d_far_e, s);
        } else {
            launch_l2p(P.d_leafdesc, P.n_far_leaves, lv, P.dX, P.dY, P.dZ, P.dL, P.k, P.p_leafmax, P.d_far, s);
        }
        mark(6);
    }
    if (parts & kSeqP2P) {
        if (field)
            launch_p2p_field(P.precision, P.n_leaves, P.n_p2p_fp32, P.d_leaf_box, bx, P.d_rng_ptr, P.d_rng_start,
                             P.d_rng_pref, P.d_leaf_src_total, P.d_floor2, P.dX, P.dY, P.dZ, P.d_q, P.k, P.d_far,
                             P.d_far_e, P.d_perm, out_mode, out, field, s);
        else
            launch_p2p(p2p_args(P, out_mode, out), P.precision, P.n_leaves, P.n_p2p_fp32, P.p2p_lowk, P.p2p_tile, s);
        mark(7);
    }
}

// one evaluate into `out` of out_type (HFMM_C64: the FP64 result lands in d_out and is
// rounded into out)
void sequence_typed(hfmm_plan& P, const void* q, int q_type, void* out, int out_type, hipStream_t s) {
    if (out_type == HFMM_C64) {
        run_sequence(P, q, q_type, P.d_out, s, nullptr);
        launch_to_c64(P.d_out, (float2*)out, P.n, s);
    } else {
        run_sequence(P, q, q_type, (double2*)out, s, nullptr);
    }
}

void evaluate_device(hfmm_plan& P, const void* q, int q_type, void* out, hipStream_t s, int out_type = HFMM_C128) {
    if (!P.use_graph) {
        sequence_typed(P, q, q_type, out, out_type, s);
        return;
    }
    if (!P.gexec || P.g_q != q || P.g_out != out || P.g_qtype != q_type || P.g_outtype != out_type) {
        if (P.gexec) {
            HIPCHK(hipGraphExecDestroy(P.gexec));
            P.gexec = nullptr;
        }
        if (P.graph) {
            HIPCHK(hipGraphDestroy(P.graph));
            P.graph = nullptr;
        }
        HIPCHK(hipStreamBeginCapture(P.stream, hipStreamCaptureModeThreadLocal));
        try {
            sequence_typed(P, q, q_type, out, out_type, P.stream);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(P.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCHK(hipStreamEndCapture(P.stream, &P.graph));
        HIPCHK(hipGraphInstantiate(&P.gexec, P.graph, nullptr, nullptr, 0));
        P.g_q = q;
        P.g_out = out;
        P.g_qtype = q_type;
        P.g_outtype = out_type;
    }
    HIPCHK(hipGraphLaunch(P.gexec, s));
}

}  // namespace

namespace {
int guard_impl(void (*fn)(void*), void* ctx) {
    try {
        fn(ctx);
        return HFMM_OK;
    } catch (const Error& e) {
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return HFMM_ERR_MEMORY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return HFMM_ERR_INTERNAL;
    } catch (...) {
        g_last_error = "unknown error";
        return HFMM_ERR_INTERNAL;
    }
}
template <typename F>
int guarded(F&& f) {
    return guard_impl([](void* c) { (*(F*)c)(); }, (void*)&f);
}
}  // namespace

static void export_tree(const Tree& t, int* level, uint64_t* code, int* count, int* pstart, int* parent,
                        uint8_t* is_leaf, int64_t* perm);
static void export_list(const Lists& ls, char which, int* pairs);

extern "C" {

void hfmm_opts_default(hfmm_opts* o) {
    std::memset(o, 0, sizeof(*o));
    o->p = 4;
    o->n_max = 16;
    o->max_depth = 0;
    o->precision = HFMM_PREC_FP64;
    o->use_graph = 1;
    o->m2l_kernel = 0;
    o->m2l_fp64_levels = -1;
    o->p2p_tile = 0;
}

int hfmm_plan_create(const double* xyz, int64_t n, double k, const hfmm_opts* o, hfmm_plan** out) {
    return guarded([&] {
        if (!out || !o || !xyz) fail(HFMM_ERR_ARG, "null argument");
        *out = nullptr;
        hfmm_plan* P = new hfmm_plan();
        try {
            build_plan(*P, xyz, n, k, *o);
        } catch (...) {
            delete P;
            throw;
        }
        *out = P;
    });
}

int hfmm_plan_info_get(const hfmm_plan* P, hfmm_plan_info* info) {
    return guarded([&] {
        if (!P || !info) fail(HFMM_ERR_ARG, "null argument");
        std::memset(info, 0, sizeof(*info));
        info->n_points = P->n;
        info->depth = P->depth;
        info->n_boxes = P->tree.n_boxes();
        info->n_leaves = P->n_leaves;
        info->n_leaves_phase64 = P->precision == HFMM_PREC_MIXED ? P->n_leaves - P->n_p2p_fp32 : 0;
        info->maxleaf = P->tree.maxleaf;
        info->n_v = (int64_t)P->lists.v_tgt.size();
        info->n_u = (int64_t)P->lists.u_tgt.size();
        info->n_w = (int64_t)P->lists.w_tgt.size();
        info->n_x = (int64_t)P->lists.x_tgt.size();
        info->n_p2p_pairs = P->n_p2p_pairs;
        for (int L = 0; L < HFMM_MAX_LEVELS; ++L) {
            info->p_level[L] = (L >= 1 && L <= P->depth) ? P->p_eff[L] : 0;
            info->p_table[L] = (L >= 1 && L <= P->depth) ? P->p_level[L] : 0;
            if (P->m2lb[L].use && P->m2lb[L].fp32) info->m2l_fp32_levels |= 1 << L;
            if (P->m2mb[L].use && P->m2mb[L].fp32) info->m2m_fp32_levels |= 1 << L;
            if (P->l2lb[L].use && P->l2lb[L].fp32) info->l2l_fp32_levels |= 1 << L;
        }
        info->n_leaves_coord32 = P->n_leaves_coord32;
        info->extent = P->tree.extent;
        for (int a = 0; a < 3; ++a) info->pmin[a] = P->tree.pmin[a];
    });
}

int hfmm_plan_export_tree(const hfmm_plan* P, int* level, uint64_t* code, int* count, int* pstart, int* parent,
                          uint8_t* is_leaf, int64_t* perm) {
    return guarded([&] {
        if (!P) fail(HFMM_ERR_ARG, "null plan");
        export_tree(P->tree, level, code, count, pstart, parent, is_leaf, perm);
    });
}

int hfmm_plan_export_list(const hfmm_plan* P, char which, int* pairs) {
    return guarded([&] {
        if (!P || !pairs) fail(HFMM_ERR_ARG, "null argument");
        export_list(P->lists, which, pairs);
    });
}

int hfmm_plan_set_degrees(hfmm_plan* P, const int* p_level) {
    return guarded([&] {
        if (!P || !p_level) fail(HFMM_ERR_ARG, "null argument");
        int eff[kMaxLevels];
        for (int L = 0; L < kMaxLevels; ++L) {
            eff[L] = P->p_level[L];
            if (L >= HFMM_MAX_LEVELS || L < 1 || L > P->depth || p_level[L] == 0) continue;
            if (p_level[L] < 2 || p_level[L] > P->p_level[L])
                fail(HFMM_ERR_ARG, "p_level entries must be 0 or in [2, table degree of the level]");
            eff[L] = p_level[L];
        }
        HIPCHK(hipStreamSynchronize(P->stream));
        std::copy(eff, eff + kMaxLevels, P->p_eff);
        ++P->degree_version;
        HIPCHK(hipMemcpy(P->d_plevel, P->p_eff, sizeof(int) * kMaxLevels, hipMemcpyHostToDevice));
        if (P->gexec) {
            HIPCHK(hipGraphExecDestroy(P->gexec));
            P->gexec = nullptr;
        }
        if (P->graph) {
            HIPCHK(hipGraphDestroy(P->graph));
            P->graph = nullptr;
        }
        P->g_q = nullptr;
    });
}

int hfmm_evaluate_typed(hfmm_plan* P, const void* q, int q_type, void* out, int out_type, int ptr_kind,
                        void* stream) {
    return guarded([&] {
        if (!P || !q || !out) fail(HFMM_ERR_ARG, "null argument");
        if (q_type < HFMM_F64 || q_type > HFMM_C64) fail(HFMM_ERR_ARG, "unknown charge type");
        if (out_type != HFMM_C128 && out_type != HFMM_C64) fail(HFMM_ERR_ARG, "unknown output type");
        hipStream_t s = stream ? (hipStream_t)stream : P->stream;
        if (ptr_kind == HFMM_HOST) {
            HIPCHK(hipMemcpyAsync(P->d_qin, q, (size_t)P->n * charge_bytes(q_type), hipMemcpyHostToDevice, s));
            if (out_type == HFMM_C64) {
                if (!P->d_out32) P->d_out32 = P->alloc<float2>(P->n);
                evaluate_device(*P, P->d_qin, q_type, P->d_out32, s, HFMM_C64);
                HIPCHK(hipMemcpyAsync(out, P->d_out32, (size_t)P->n * 8, hipMemcpyDeviceToHost, s));
            } else {
                evaluate_device(*P, P->d_qin, q_type, P->d_out, s);
                HIPCHK(hipMemcpyAsync(out, P->d_out, (size_t)P->n * 16, hipMemcpyDeviceToHost, s));
            }
        } else if (ptr_kind == HFMM_DEVICE) {
            evaluate_device(*P, q, q_type, out, s, out_type);
        } else {
            fail(HFMM_ERR_ARG, "unknown pointer kind");
        }
        HIPCHK(hipStreamSynchronize(s));
    });
}

int hfmm_evaluate_field(hfmm_plan* P, const void* q, int q_type, void* phi, void* field, int ptr_kind,
                        void* stream) {
    return guarded([&] {
        if (!P || !q || !phi || !field) fail(HFMM_ERR_ARG, "null argument");
        if (q_type < HFMM_F64 || q_type > HFMM_C64) fail(HFMM_ERR_ARG, "unknown charge type");
        if (P->p_leafmax + 1 > kMaxP) fail(HFMM_ERR_ARG, "the field needs leaf degrees of at most 39");
        hipStream_t s = stream ? (hipStream_t)stream : P->stream;
        if (!P->d_far_e) P->d_far_e = P->alloc<double2>(3 * P->n);
        if (ptr_kind == HFMM_HOST) {
            if (!P->d_field_out) P->d_field_out = P->alloc<double2>(3 * P->n);
            HIPCHK(hipMemcpyAsync(P->d_qin, q, (size_t)P->n * charge_bytes(q_type), hipMemcpyHostToDevice, s));
            run_sequence(*P, P->d_qin, q_type, P->d_out, s, nullptr, nullptr, P->d_field_out);
            HIPCHK(hipMemcpyAsync(phi, P->d_out, (size_t)P->n * 16, hipMemcpyDeviceToHost, s));
            HIPCHK(hipMemcpyAsync(field, P->d_field_out, (size_t)P->n * 48, hipMemcpyDeviceToHost, s));
        } else if (ptr_kind == HFMM_DEVICE) {
            run_sequence(*P, q, q_type, (double2*)phi, s, nullptr, nullptr, (double2*)field);
        } else {
            fail(HFMM_ERR_ARG, "unknown pointer kind");
        }
        HIPCHK(hipStreamSynchronize(s));
    });
}

int hfmm_evaluate(hfmm_plan* P, const void* q, int q_type, void* out, int ptr_kind, void* stream) {
    return hfmm_evaluate_typed(P, q, q_type, out, HFMM_C128, ptr_kind, stream);
}

int hfmm_benchmark(hfmm_plan* P, const void* q_dev, int q_type, void* out_dev, int warmup, int reps,
                   double* median_ms, double* phase_ms) {
    return guarded([&] {
        if (!P || !q_dev || !out_dev || reps < 1) fail(HFMM_ERR_ARG, "bad argument");
        hipStream_t s = P->stream;
        for (int i = 0; i < warmup; ++i) evaluate_device(*P, q_dev, q_type, out_dev, s);
        HIPCHK(hipStreamSynchronize(s));
        std::vector<hipEvent_t> ev(2 * reps);
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        for (int i = 0; i < reps; ++i) {
            HIPCHK(hipEventRecord(ev[2 * i], s));
            evaluate_device(*P, q_dev, q_type, out_dev, s);
            HIPCHK(hipEventRecord(ev[2 * i + 1], s));
        }
        HIPCHK(hipStreamSynchronize(s));
        std::vector<double> ms(reps);
        for (int i = 0; i < reps; ++i) {
            float f = 0;
            HIPCHK(hipEventElapsedTime(&f, ev[2 * i], ev[2 * i + 1]));
            ms[i] = f;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        std::sort(ms.begin(), ms.end());
        if (median_ms) *median_ms = ms[reps / 2];
        if (phase_ms) {
            hipEvent_t pe[8];
            for (auto& e : pe) HIPCHK(hipEventCreate(&e));
            for (int i = 0; i < 7; ++i) phase_ms[i] = 0.0;
            for (int r = 0; r < reps; ++r) {
                run_sequence(*P, q_dev, q_type, (double2*)out_dev, s, pe);
                HIPCHK(hipStreamSynchronize(s));
                for (int i = 0; i < 7; ++i) {
                    float f = 0;
                    HIPCHK(hipEventElapsedTime(&f, pe[i], pe[i + 1]));
                    phase_ms[i] += f / reps;
                }
            }
            for (auto& e : pe) (void)hipEventDestroy(e);
        }
    });
}

int hfmm_benchmark_m2l_levels(hfmm_plan* P, const void* q_dev, int q_type, void* out_dev, int reps,
                              double* level_ms) {
    return guarded([&] {
        if (!P || !q_dev || !out_dev || !level_ms || reps < 1) fail(HFMM_ERR_ARG, "bad argument");
        hipStream_t s = P->stream;
        std::vector<hipEvent_t> ev(2 * kMaxLevels);
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        for (int L = 0; L < HFMM_MAX_LEVELS; ++L) level_ms[L] = 0.0;
        for (int r = 0; r < reps; ++r) {
            run_sequence(*P, q_dev, q_type, (double2*)out_dev, s, nullptr, ev.data());
            HIPCHK(hipStreamSynchronize(s));
            for (int L = 2; L <= P->depth && L < HFMM_MAX_LEVELS; ++L) {
                if (!P->m2l[L].n_targets) continue;
                float f = 0;
                HIPCHK(hipEventElapsedTime(&f, ev[2 * L], ev[2 * L + 1]));
                level_ms[L] += f / reps;
            }
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
    });
}

int hfmm_plan_destroy(hfmm_plan* P) {
    return guarded([&] { delete P; });
}

int hfmm_direct(const double* xyz, int64_t n, const double* q, double k, const int64_t* targets, int64_t nt,
                double* out) {
    return guarded([&] {
        if (!xyz || !q || !out || (nt > 0 && !targets) || n <= 0 || nt < 0) fail(HFMM_ERR_ARG, "bad argument");
        if (!std::isfinite(k)) fail(HFMM_ERR_ARG, "k must be a finite real number");
        check_device();
        std::vector<double> X(n), Y(n), Z(n), tx(nt), ty(nt), tz(nt);
        for (int64_t i = 0; i < n; ++i) {
            X[i] = xyz[3 * i];
            Y[i] = xyz[3 * i + 1];
            Z[i] = xyz[3 * i + 2];
        }
        for (int64_t i = 0; i < nt; ++i) {
            if (targets[i] < 0 || targets[i] >= n) fail(HFMM_ERR_ARG, "target index out of range");
            tx[i] = X[targets[i]];
            ty[i] = Y[targets[i]];
            tz[i] = Z[targets[i]];
        }
        hfmm_plan tmp;   // owns the temporaries
        double* dX = tmp.upload(X);
        double* dY = tmp.upload(Y);
        double* dZ = tmp.upload(Z);
        double* dtx = tmp.upload(tx);
        double* dty = tmp.upload(ty);
        double* dtz = tmp.upload(tz);
        double2* dq = tmp.alloc<double2>(n);
        HIPCHK(hipMemcpy(dq, q, (size_t)n * 16, hipMemcpyHostToDevice));
        double2* dout = tmp.alloc<double2>(nt);
        launch_direct(dX, dY, dZ, dq, n, k, dtx, dty, dtz, nt, dout, nullptr);
        HIPCHK(hipDeviceSynchronize());
        if (nt) HIPCHK(hipMemcpy(out, dout, (size_t)nt * 16, hipMemcpyDeviceToHost));
    });
}

struct hfmm_tree {
    Tree tree;
    Lists lists;
};

int hfmm_tree_create(const double* xyz, int64_t n, int n_max, int max_depth, hfmm_tree** out) {
    return guarded([&] {
        if (!out || !xyz) fail(HFMM_ERR_ARG, "null argument");
        *out = nullptr;
        if (n <= 0 || n > (int64_t)1 << 30) fail(HFMM_ERR_ARG, "positions must be a nonempty finite (N,3) array");
        for (int64_t i = 0; i < 3 * n; ++i)
            if (!std::isfinite(xyz[i])) fail(HFMM_ERR_ARG, "positions must be a nonempty finite (N,3) array");
        if (n_max < 1) fail(HFMM_ERR_ARG, "p must be an integer >=2; N_max must be a positive integer");
        hfmm_tree* T = new hfmm_tree();
        try {
            T->tree = build_tree(xyz, n, n_max, max_depth);
            T->lists = build_lists(T->tree);
        } catch (...) {
            delete T;
            throw;
        }
        *out = T;
    });
}

int hfmm_tree_create_device(const double* xyz, int64_t n, int n_max, int max_depth, hfmm_tree** out) {
    return guarded([&] {
        if (!out || !xyz) fail(HFMM_ERR_ARG, "null argument");
        *out = nullptr;
        if (n <= 0 || n > (int64_t)1 << 30) fail(HFMM_ERR_ARG, "positions must be a nonempty finite (N,3) array");
        for (int64_t i = 0; i < 3 * n; ++i)
            if (!std::isfinite(xyz[i])) fail(HFMM_ERR_ARG, "positions must be a nonempty finite (N,3) array");
        if (n_max < 1) fail(HFMM_ERR_ARG, "p must be an integer >=2; N_max must be a positive integer");
        check_device();
        std::unique_ptr<hfmm_tree> T(new hfmm_tree());
        build_tree_device(xyz, n, n_max, max_depth, T->tree, T->lists);
        *out = T.release();
    });
}

int hfmm_tree_info_get(const hfmm_tree* T, hfmm_plan_info* info) {
    return guarded([&] {
        if (!T || !info) fail(HFMM_ERR_ARG, "null argument");
        std::memset(info, 0, sizeof(*info));
        info->n_points = T->tree.n_points;
        info->depth = T->tree.depth;
        info->n_boxes = T->tree.n_boxes();
        int nl = 0;
        for (auto f : T->tree.is_leaf) nl += f;
        info->n_leaves = nl;
        info->maxleaf = T->tree.maxleaf;
        info->n_v = (int64_t)T->lists.v_tgt.size();
        info->n_u = (int64_t)T->lists.u_tgt.size();
        info->n_w = (int64_t)T->lists.w_tgt.size();
        info->n_x = (int64_t)T->lists.x_tgt.size();
        info->extent = T->tree.extent;
        for (int a = 0; a < 3; ++a) info->pmin[a] = T->tree.pmin[a];
    });
}

static void export_tree(const Tree& t, int* level, uint64_t* code, int* count, int* pstart, int* parent,
                        uint8_t* is_leaf, int64_t* perm) {
    if (level) std::copy(t.level.begin(), t.level.end(), level);
    if (code) std::copy(t.code.begin(), t.code.end(), code);
    if (count) std::copy(t.count.begin(), t.count.end(), count);
    if (pstart) std::copy(t.pstart.begin(), t.pstart.end(), pstart);
    if (parent) std::copy(t.parent.begin(), t.parent.end(), parent);
    if (is_leaf) std::copy(t.is_leaf.begin(), t.is_leaf.end(), is_leaf);
    if (perm) std::copy(t.perm.begin(), t.perm.end(), perm);
}

static void export_list(const Lists& ls, char which, int* pairs) {
    const std::vector<int>*a, *b;
    switch (which) {
        case 'V': a = &ls.v_tgt; b = &ls.v_src; break;
        case 'U': a = &ls.u_tgt; b = &ls.u_src; break;
        case 'W': a = &ls.w_tgt; b = &ls.w_src; break;
        case 'X': a = &ls.x_tgt; b = &ls.x_src; break;
        default: fail(HFMM_ERR_ARG, "list must be one of V, U, W, X");
    }
    for (size_t i = 0; i < a->size(); ++i) {
        pairs[2 * i] = (*a)[i];
        pairs[2 * i + 1] = (*b)[i];
    }
}

int hfmm_tree_export(const hfmm_tree* T, int* level, uint64_t* code, int* count, int* pstart, int* parent,
                     uint8_t* is_leaf, int64_t* perm) {
    return guarded([&] {
        if (!T) fail(HFMM_ERR_ARG, "null tree");
        export_tree(T->tree, level, code, count, pstart, parent, is_leaf, perm);
    });
}

int hfmm_tree_export_list(const hfmm_tree* T, char which, int* pairs) {
    return guarded([&] {
        if (!T || !pairs) fail(HFMM_ERR_ARG, "null argument");
        export_list(T->lists, which, pairs);
    });
}

int hfmm_tree_destroy(hfmm_tree* T) {
    return guarded([&] { delete T; });
}

int hfmm_host_jhat(double x, int p, double* out) {
    return guarded([&] {
        if (!out || p < 0 || p > 4 * kMaxP) fail(HFMM_ERR_ARG, "bad argument");
        jhat_host(x, p, out);
    });
}

int hfmm_host_hhat(double x, int p, double* re, double* im) {
    return guarded([&] {
        if (!re || !im || p < 1 || p > 4 * kMaxP) fail(HFMM_ERR_ARG, "bad argument");
        hhat_host(x, p, re, im);
    });
}

int hfmm_host_rotation(double theta, int p, double* out) {
    return guarded([&] {
        if (!out || p < 0 || p > kMaxP) fail(HFMM_ERR_ARG, "bad argument");
        rotation_table(theta, p, out);
    });
}

int hfmm_host_coax(int kind, double k, double t, double rs, double rt, int ps, int pt, double* out) {
    return guarded([&] {
        if (!out || kind < 0 || kind > 2 || ps < 0 || pt < 0 || ps > kMaxP || pt > kMaxP || !(t > 0) ||
            !(rs > 0) || !(rt > 0) || !(k >= 0))
            fail(HFMM_ERR_ARG, "bad argument");
        const Gaunt G = make_gaunt(std::max(ps, pt));
        coax_table(kind, G, k, t, rs, rt, ps, pt, out);
    });
}

const char* hfmm_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

// ---------------------------------------------------------------------------------
// Batches of independent problems: one plan per problem; an evaluate forks the
// problems over a few streams (balanced by point count) and joins them, all inside
// one captured hipGraph, so that small problems run side by side on the GPU.
// ---------------------------------------------------------------------------------
struct hfmm_batch {
    std::vector<hfmm_plan*> plans;
    std::vector<std::vector<int>> branch;   // problems per stream, largest first
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> join, join2;
    hipEvent_t fork = nullptr, fork2 = nullptr;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<const void*> g_q;
    std::vector<void*> g_out;
    std::vector<unsigned> g_ver;            // degree versions of the plans at capture
    int g_qtype = -1;
    // persistent launches over all problems (rebuilt with the graph): the block M2L, one
    // launch per kernel instance, and the near field
    struct M2LInstance {
        int p = 0, tb = 0, g = 0, fp32 = 0, grid = 0;
        size_t shm = 0;
        WorkQueues q{};
    };
    std::vector<M2LInstance> m2l_inst;
    std::vector<const M2LBlockArgs*> m2l_args;   // device descriptors per instance
    const P2PBatchArgs* d_p2p_args = nullptr;
    WorkQueues p2p_q[4]{};                 // per near-field leaf class (kP2PClass*)
    int p2p_grid[4] = {0, 0, 0, 0};
    int* d_counters = nullptr;
    int n_counters = 0;
    std::vector<void*> pallocs;             // device buffers of the persistent launches
    void free_persistent() {
        for (void* p : pallocs) (void)hipFree(p);
        pallocs.clear();
        m2l_inst.clear();
        m2l_args.clear();
        d_p2p_args = nullptr;
        d_counters = nullptr;
    }
    template <typename T>
    T* upload(const std::vector<T>& v) {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(1, v.size()) * sizeof(T)) != hipSuccess)
            fail(HFMM_ERR_MEMORY, "hipMalloc failed (batch work lists)");
        pallocs.push_back(p);
        if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return (T*)p;
    }
    ~hfmm_batch() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        free_persistent();
        for (auto e : join) (void)hipEventDestroy(e);
        for (auto e : join2) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (fork2) (void)hipEventDestroy(fork2);
        for (auto st : streams) (void)hipStreamDestroy(st);
        if (stream) (void)hipStreamDestroy(stream);
        for (auto P : plans) delete P;
    }
};

namespace {

constexpr int kBatchStreams = 4;

// kernel instance of the block M2L (the table-degree range selects KT and MAXU)
int m2l_degree_range(int p) { return p <= 7 ? 0 : p <= 11 ? 1 : p <= 15 ? 2 : p <= 17 ? 3 : p <= 19 ? 4 : 5; }

// queues of a work list: kXcds contiguous ranges of equal length
WorkQueues make_queues(hfmm_batch& B, const std::vector<int2>& work, int* counters) {
    std::vector<int> start(kXcds + 1);
    for (int x = 0; x <= kXcds; ++x) start[x] = (int)((int64_t)work.size() * x / kXcds);
    WorkQueues q;
    q.work = B.upload(work);
    q.start = B.upload(start);
    q.count = counters;
    return q;
}

// Device descriptors and work lists of the persistent launches for the current degrees
// and pointers: the block M2L levels of every problem grouped by kernel instance, and the
// leaves of every problem for the near field.
void build_persistent(hfmm_batch& B, void* const* out, int q_type) {
    B.free_persistent();
    const int np = (int)B.plans.size();
    std::map<std::array<int, 4>, std::vector<M2LBlockArgs>> groups;
    std::map<std::array<int, 4>, int> pmax;
    for (int i = 0; i < np; ++i) {
        hfmm_plan& P = *B.plans[i];
        for (int L = 2; L <= P.depth; ++L) {
            const M2LBlockLevel& lb = P.m2lb[L];
            if (!P.m2l[L].n_targets || !lb.use || lb.n_items <= 0) continue;
            const int pk = std::max(lb.ps, lb.pt);
            const std::array<int, 4> key{m2l_degree_range(pk), lb.tb, lb.grp, lb.fp32 ? 1 : 0};
            groups[key].push_back(block_args(P, lb, lb.fp32, lb.fp32 ? (const void*)P.dM32 : (const void*)P.dM, P.dL, 0));
            pmax[key] = std::max(pmax[key], pk);
        }
    }
    B.n_counters = 16 * kXcds * ((int)groups.size() + 4);
    std::vector<int> zeros(B.n_counters, 0);
    B.d_counters = B.upload(zeros);
    int qi = 0;
    for (auto& kv : groups) {
        hfmm_batch::M2LInstance in;
        in.p = pmax[kv.first];
        in.tb = kv.first[1];
        in.g = kv.first[2];
        in.fp32 = kv.first[3];
        in.shm = m2l_block_lds_bytes(in.p, in.tb, in.g, in.fp32 ? 4 : 8);
        std::vector<int2> work;
        for (int j = 0; j < (int)kv.second.size(); ++j)
            for (int t = 0; t < kv.second[j].n_items; ++t) work.push_back(make_int2(j, t));
        const M2LBlockArgs* d_args = B.upload(kv.second);
        in.q = make_queues(B, work, B.d_counters + 16 * kXcds * qi++);
        in.grid = m2l_block_persistent_grid(in.p, in.tb, in.g, in.fp32, in.shm);
        B.m2l_inst.push_back(in);
        B.m2l_args.push_back(d_args);
    }
    std::vector<P2PBatchArgs> pa(np);
    std::vector<int2> work[4];
    for (int i = 0; i < np; ++i) {
        hfmm_plan& P = *B.plans[i];
        pa[i].a = p2p_scaled(p2p_args(P, out_mode_of(P, q_type), (double2*)out[i]), P.precision, P.p2p_lowk);
        for (int l = 0; l < P.n_leaves; ++l) {
            const int cls = P.precision == HFMM_PREC_FP64 ? kP2PClassFp64
                            : P.p2p_lowk                  ? kP2PClassLowK
                            : l < P.n_p2p_fp32            ? kP2PClassPhase32
                                                          : kP2PClassPhase64;
            work[cls].push_back(make_int2(i, l));
        }
    }
    B.d_p2p_args = B.upload(pa);
    for (int c = 0; c < 4; ++c) {
        B.p2p_q[c] = make_queues(B, work[c], B.d_counters + 16 * kXcds * (qi + c));
        B.p2p_grid[c] = work[c].empty() ? 0 : p2p_persistent_grid(c);
    }
}

// ev (optional, 5 events): brackets of the upward part, the persistent M2L, the downward
// part and the persistent near field
void batch_sequence(hfmm_batch& B, const void* const* q, int q_type, void* const* out, hipStream_t s,
                    hipEvent_t* ev = nullptr) {
    auto mark = [&](int i) {
        if (ev) HIPCHK(hipEventRecord(ev[i], s));
    };
    mark(0);
    launch_clear_ints(B.d_counters, B.n_counters, s);
    // upward passes and the generic-kernel M2L levels, problems forked over the streams
    HIPCHK(hipEventRecord(B.fork, s));
    for (size_t b = 0; b < B.streams.size(); ++b) {
        HIPCHK(hipStreamWaitEvent(B.streams[b], B.fork, 0));
        for (int i : B.branch[b])
            run_sequence(*B.plans[i], q[i], q_type, (double2*)out[i], B.streams[b], nullptr, nullptr, nullptr,
                         kSeqUp | kSeqM2LGeneric);
        HIPCHK(hipEventRecord(B.join[b], B.streams[b]));
        HIPCHK(hipStreamWaitEvent(s, B.join[b], 0));
    }
    mark(1);
    // the block M2L of every problem and level: persistent launches
    for (size_t j = 0; j < B.m2l_inst.size(); ++j) {
        const auto& in = B.m2l_inst[j];
        launch_m2l_block_persistent(B.m2l_args[j], in.q, in.grid, in.p, in.tb, in.g, in.fp32, in.shm, s);
    }
    mark(2);
    // downward passes
    HIPCHK(hipEventRecord(B.fork2, s));
    for (size_t b = 0; b < B.streams.size(); ++b) {
        HIPCHK(hipStreamWaitEvent(B.streams[b], B.fork2, 0));
        for (int i : B.branch[b])
            run_sequence(*B.plans[i], q[i], q_type, (double2*)out[i], B.streams[b], nullptr, nullptr, nullptr, kSeqDown);
        HIPCHK(hipEventRecord(B.join2[b], B.streams[b]));
        HIPCHK(hipStreamWaitEvent(s, B.join2[b], 0));
    }
    mark(3);
    // the near field of every problem: one persistent launch per leaf class
    for (int c = 0; c < 4; ++c) launch_p2p_persistent(c, B.d_p2p_args, B.p2p_q[c], B.p2p_grid[c], s);
    mark(4);
}

void batch_evaluate_device(hfmm_batch& B, const void* const* q, int q_type, void* const* out, hipStream_t s) {
    const size_t np = B.plans.size();
    bool same = B.gexec && B.g_qtype == q_type;
    for (size_t i = 0; same && i < np; ++i)
        same = B.g_q[i] == q[i] && B.g_out[i] == out[i] && B.g_ver[i] == B.plans[i]->degree_version;
    if (!same) {
        if (B.gexec) HIPCHK(hipGraphExecDestroy(B.gexec));
        if (B.graph) HIPCHK(hipGraphDestroy(B.graph));
        B.gexec = nullptr;
        B.graph = nullptr;
        build_persistent(B, out, q_type);
        HIPCHK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
        try {
            batch_sequence(B, q, q_type, out, B.stream);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(B.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCHK(hipStreamEndCapture(B.stream, &B.graph));
        HIPCHK(hipGraphInstantiate(&B.gexec, B.graph, nullptr, nullptr, 0));
        B.g_q.assign(q, q + np);
        B.g_out.assign(out, out + np);
        B.g_ver.resize(np);
        for (size_t i = 0; i < np; ++i) B.g_ver[i] = B.plans[i]->degree_version;
        B.g_qtype = q_type;
    }
    HIPCHK(hipGraphLaunch(B.gexec, s));
}

}  // namespace

extern "C" {

int hfmm_batch_create(int n_problems, const double* const* xyz, const int64_t* n, const double* k,
                      const hfmm_opts* opts, hfmm_batch** out) {
    return guarded([&] {
        if (!out || !xyz || !n || !k || !opts || n_problems < 1) fail(HFMM_ERR_ARG, "bad argument");
        *out = nullptr;
        std::unique_ptr<hfmm_batch> B(new hfmm_batch());
        for (int i = 0; i < n_problems; ++i) {
            if (!xyz[i]) fail(HFMM_ERR_ARG, "null positions of problem " + std::to_string(i));
            hfmm_plan* P = new hfmm_plan();
            try {
                build_plan(*P, xyz[i], n[i], k[i], opts[i]);
            } catch (const Error& e) {
                delete P;
                fail(e.code, "problem " + std::to_string(i) + ": " + e.msg);
            }
            B->plans.push_back(P);
        }
        const int ns = std::min(kBatchStreams, n_problems);
        B->branch.resize(ns);
        std::vector<int> order(n_problems);
        for (int i = 0; i < n_problems; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n[a] > n[b]; });
        std::vector<int64_t> load(ns, 0);
        for (int i : order) {
            const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            B->branch[b].push_back(i);
            load[b] += n[i];
        }
        HIPCHK(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
        B->streams.resize(ns);
        B->join.resize(ns);
        B->join2.resize(ns);
        for (int b = 0; b < ns; ++b) {
            HIPCHK(hipStreamCreateWithFlags(&B->streams[b], hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&B->join[b], hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&B->join2[b], hipEventDisableTiming));
        }
        HIPCHK(hipEventCreateWithFlags(&B->fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&B->fork2, hipEventDisableTiming));
        *out = B.release();
    });
}

int hfmm_batch_plan(hfmm_batch* B, int i, hfmm_plan** plan) {
    return guarded([&] {
        if (!B || !plan || i < 0 || i >= (int)B->plans.size()) fail(HFMM_ERR_ARG, "bad argument");
        *plan = B->plans[i];
    });
}

int hfmm_batch_evaluate(hfmm_batch* B, const void* const* q, int q_type, void* const* out, int ptr_kind,
                        void* stream) {
    return guarded([&] {
        if (!B || !q || !out) fail(HFMM_ERR_ARG, "null argument");
        if (q_type < HFMM_F64 || q_type > HFMM_C64) fail(HFMM_ERR_ARG, "unknown charge type");
        const size_t np = B->plans.size();
        for (size_t i = 0; i < np; ++i)
            if (!q[i] || !out[i]) fail(HFMM_ERR_ARG, "null charges or output of problem " + std::to_string(i));
        hipStream_t s = stream ? (hipStream_t)stream : B->stream;
        if (ptr_kind == HFMM_HOST) {
            std::vector<const void*> qd(np);
            std::vector<void*> od(np);
            for (size_t i = 0; i < np; ++i) {
                hfmm_plan& P = *B->plans[i];
                HIPCHK(hipMemcpyAsync(P.d_qin, q[i], (size_t)P.n * charge_bytes(q_type), hipMemcpyHostToDevice, s));
                qd[i] = P.d_qin;
                od[i] = P.d_out;
            }
            batch_evaluate_device(*B, qd.data(), q_type, od.data(), s);
            for (size_t i = 0; i < np; ++i)
                HIPCHK(hipMemcpyAsync(out[i], B->plans[i]->d_out, (size_t)B->plans[i]->n * 16,
                                      hipMemcpyDeviceToHost, s));
        } else if (ptr_kind == HFMM_DEVICE) {
            batch_evaluate_device(*B, q, q_type, out, s);
        } else {
            fail(HFMM_ERR_ARG, "unknown pointer kind");
        }
        HIPCHK(hipStreamSynchronize(s));
    });
}

int hfmm_batch_benchmark(hfmm_batch* B, const void* const* q_dev, int q_type, void* const* out_dev, int warmup,
                         int reps, double* median_ms) {
    return guarded([&] {
        if (!B || !q_dev || !out_dev || reps < 1 || !median_ms) fail(HFMM_ERR_ARG, "bad argument");
        hipStream_t s = B->stream;
        for (int i = 0; i < warmup; ++i) batch_evaluate_device(*B, q_dev, q_type, out_dev, s);
        HIPCHK(hipStreamSynchronize(s));
        std::vector<hipEvent_t> ev(2 * reps);
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        for (int i = 0; i < reps; ++i) {
            HIPCHK(hipEventRecord(ev[2 * i], s));
            batch_evaluate_device(*B, q_dev, q_type, out_dev, s);
            HIPCHK(hipEventRecord(ev[2 * i + 1], s));
        }
        HIPCHK(hipStreamSynchronize(s));
        std::vector<double> ms(reps);
        for (int i = 0; i < reps; ++i) {
            float f = 0;
            HIPCHK(hipEventElapsedTime(&f, ev[2 * i], ev[2 * i + 1]));
            ms[i] = f;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        std::sort(ms.begin(), ms.end());
        *median_ms = ms[reps / 2];
    });
}

int hfmm_batch_benchmark_phases(hfmm_batch* B, const void* const* q_dev, int q_type, void* const* out_dev, int reps,
                                double* phase_ms) {
    return guarded([&] {
        if (!B || !q_dev || !out_dev || reps < 1 || !phase_ms) fail(HFMM_ERR_ARG, "bad argument");
        hipStream_t s = B->stream;
        batch_evaluate_device(*B, q_dev, q_type, out_dev, s);   // descriptors for these pointers
        HIPCHK(hipStreamSynchronize(s));
        hipEvent_t ev[5];
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        for (int i = 0; i < 4; ++i) phase_ms[i] = 0.0;
        for (int r = 0; r < reps; ++r) {
            batch_sequence(*B, q_dev, q_type, out_dev, s, ev);
            HIPCHK(hipStreamSynchronize(s));
            for (int i = 0; i < 4; ++i) {
                float f = 0;
                HIPCHK(hipEventElapsedTime(&f, ev[i], ev[i + 1]));
                phase_ms[i] += f / reps;
            }
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
    });
}

int hfmm_batch_destroy(hfmm_batch* B) {
    return guarded([&] { delete B; });
}

}  // extern "C"
