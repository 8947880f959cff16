// Octree and interaction lists built on the GPU, with the semantics, canonical order and
// arrays of the host builder (tree.cpp, which follows jaxfmm/trees/cubic.py:316-695 and
// jaxfmm/trees/lists.py:619-931); the results are identical, element for element.
//
// Points: Morton codes at the finest level and a stable radix sort of (code, index)
// (rocPRIM device primitives: radix sort, run-length encoding, reduce-by-key, select).
// Boxes: run-length encoding of the sorted codes gives the finest boxes; then, level by
// level from the finest, parents by reduce-by-key on code >> 3, the merge test (count at
// most N_max and the 2:1 balance check over the 26 neighbours) one thread per parent,
// and compaction of the children of merged parents.  The automatic depth rule repeats
// the build one level deeper while the largest leaf holds more than 2 N_max points.
// Lists: one thread per box counts its V / U / W / X entries, a scan gives offsets, the
// same thread fills them, and a radix sort on (target, source) fixes the order.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce_by_key.hpp>
#include <rocprim/device/device_run_length_encode.hpp>
#include <rocprim/device/device_select.hpp>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "common.h"
#include "morton.h"

namespace hfmm {
namespace {

void tchk(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(-3, std::string("GPU tree build: ") + what + ": " + hipGetErrorString(e));
}

// device buffer owned by the build
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { resize(count); }
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    DBuf& operator=(DBuf&& o) noexcept {
        std::swap(p, o.p);
        std::swap(n, o.n);
        return *this;
    }
    void resize(size_t count) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = count;
        if (count) tchk(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
    }
    std::vector<T> host(size_t count) const {
        std::vector<T> v(count);
        if (count) tchk(hipMemcpy(v.data(), p, count * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
        return v;
    }
    void upload(const std::vector<T>& v) {
        resize(v.size());
        if (!v.empty()) tchk(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
    }
};

// temporary storage of the hipCUB calls
struct Temp {
    DBuf<char> buf;
    void* get(size_t bytes) {
        if (bytes > buf.n) buf.resize(bytes);
        return buf.p;
    }
};

template <typename T>
T read1(const T* d) {
    T v;
    tchk(hipMemcpy(&v, d, sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
    return v;
}

unsigned blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

int bits_for(uint64_t v) {
    int b = 1;
    while (b < 64 && (v >> b)) ++b;
    return b;
}

__device__ int lower(const uint64_t* a, int n, uint64_t c) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__device__ int find_code(const uint64_t* a, int n, uint64_t c) {
    const int i = lower(a, n, c);
    return (i < n && a[i] == c) ? i : -1;
}

// Morton code of every point at the finest level (cubic.py:390-391 arithmetic order)
__global__ void codes_kernel(const double* xyz, int64_t n, double p0, double p1, double p2, double extent,
                             double grid, uint64_t gmax, uint64_t* codes, int* idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double pm[3] = {p0, p1, p2};
    uint64_t g[3];
    for (int a = 0; a < 3; ++a) {
        const double s = (xyz[3 * i + a] - pm[a]) / extent * grid;
        const uint64_t u = s > 0 ? (uint64_t)s : 0;
        g[a] = u < gmax ? u : gmax;
    }
    codes[i] = morton_encode(g[0], g[1], g[2]);
    idx[i] = (int)i;
}

__global__ void fill_u8_kernel(uint8_t* a, int n, uint8_t v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void shift_kernel(const uint64_t* in, int n, uint64_t* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i] >> 3;
}

// child octant c of the neighbour in direction (dx, dy, dz) touches the centre box
__device__ bool adj_child(int dx, int dy, int dz, int c) {
    const int d[3] = {dx, dy, dz};
    for (int a = 0; a < 3; ++a) {
        const int b = (c >> a) & 1;
        if (d[a] == -1 && b != 1) return false;
        if (d[a] == 1 && b != 0) return false;
    }
    return true;
}

__device__ bool masked_internal(const uint64_t* code, const uint8_t* leaf, int n, uint64_t pc, int dx, int dy,
                                int dz) {
    for (int c = 0; c < 8; ++c) {
        if (!adj_child(dx, dy, dz, c)) continue;
        const int i = find_code(code, n, pc * 8 + (uint64_t)c);
        if (i >= 0 && !leaf[i]) return true;
    }
    return false;
}

// merge decision of every parent box of level Lp (children at Lp + 1 = cur)
__global__ void merge_kernel(const uint64_t* pcode, const int* pcount, int npar, const uint64_t* ccode,
                             const uint8_t* cleaf, int ncur, const uint64_t* fcode, const uint8_t* fleaf, int nfine,
                             uint64_t last_code, int Lp, int check, int n_max, uint8_t* merge) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= npar) return;
    bool ok = pcount[b] <= n_max;
    if (ok && check) {
        int gx, gy, gz;
        morton_decode(pcode[b], gx, gy, gz);
        const int64_t size = 1ll << Lp;
        for (int d = 0; d < 26 && ok; ++d) {
            int dx, dy, dz;
            neighbour_offset(d, dx, dy, dz);
            const int64_t nx = gx + dx, ny = gy + dy, nz = gz + dz;
            if (nx < 0 || ny < 0 || nz < 0 || nx >= size || ny >= size || nz >= size) continue;
            const uint64_t nc = morton_encode(nx, ny, nz);
            if (find_code(pcode, npar, nc) >= 0) {
                if (masked_internal(ccode, cleaf, ncur, nc, dx, dy, dz)) ok = false;
            } else if (nc == 0) {
                // an absent code-0 neighbour is looked up as the last box of the child
                // level; its children (one level finer) are tested (cubic.py:466-476)
                if (masked_internal(fcode, fleaf, nfine, last_code, dx, dy, dz)) ok = false;
            }
        }
    }
    merge[b] = ok ? 1 : 0;
}

__global__ void keep_kernel(const uint64_t* ccode, int ncur, const uint64_t* pcode, int npar, const uint8_t* merge,
                            uint8_t* keep) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncur) return;
    keep[i] = merge[lower(pcode, npar, ccode[i] >> 3)] ? 0 : 1;
}

// first sorted point of every box, and its parent
__global__ void box_kernel(const uint64_t* code, const int* level, const int* ls, int depth, int nb,
                           const uint64_t* scodes, int64_t n, int* pstart, int* parent) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int L = level[b];
    const uint64_t first = code[b] << (3 * (depth - L));
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (scodes[mid] < first)
            lo = mid + 1;
        else
            hi = mid;
    }
    pstart[b] = (int)lo;
    int par = -1;
    if (L > 1) {
        const int s = ls[L - 1], e = ls[L];
        const int i = find_code(code + s, e - s, code[b] >> 3);
        par = i < 0 ? -1 : s + i;
    }
    parent[b] = par;
}

__device__ int find_box(const uint64_t* code, const int* ls, int depth, int lvl, uint64_t c) {
    if (lvl < 1 || lvl > depth) return -1;
    const int s = ls[lvl], e = ls[lvl + 1];
    const int i = find_code(code + s, e - s, c);
    return i < 0 ? -1 : s + i;
}

__device__ inline bool adj_coarse(int b, int c) { return (b <= 2 * c + 2) && (2 * c <= b + 1); }
__device__ inline bool adj_fine(int b, int f) { return (2 * b <= f + 1) && (f <= 2 * b + 2); }

// V / U / W / X of box b (lists.py:619-802 semantics, as tree.cpp).  Counting pass when
// out is null, else entries (target << 32 | source) written from off[list * nb + b].
__global__ void lists_kernel(const uint64_t* code, const int* level, const uint8_t* is_leaf, const int* ls,
                             int depth, int nb, int* cnt, const int64_t* off, uint64_t* outV, uint64_t* outU,
                             uint64_t* outW, uint64_t* outX) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const bool fill = outV != nullptr;
    int64_t pos[4] = {0, 0, 0, 0};
    if (fill)
        for (int w = 0; w < 4; ++w) pos[w] = off[(int64_t)w * nb + b];
    int count[4] = {0, 0, 0, 0};
    uint64_t* outs[4] = {outV, outU, outW, outX};
    auto emit = [&](int w, int src) {
        if (fill) outs[w][pos[w]++] = ((uint64_t)(uint32_t)b << 32) | (uint32_t)src;
        ++count[w];
    };
    const int L = level[b];
    const uint64_t c = code[b];
    int bx, by, bz;
    morton_decode(c, bx, by, bz);
    const bool leaf = is_leaf[b] != 0;
    const int px = bx >> 1, py = by >> 1, pz = bz >> 1;
    const int64_t psize = 1ll << (L - 1), size = 1ll << L;
    for (int d = 0; d < 26; ++d) {
        int dx, dy, dz;
        neighbour_offset(d, dx, dy, dz);
        const int64_t nx = px + dx, ny = py + dy, nz = pz + dz;
        if (nx < 0 || ny < 0 || nz < 0 || nx >= psize || ny >= psize || nz >= psize) continue;
        const uint64_t nc = morton_encode(nx, ny, nz);
        for (int ch = 0; ch < 8; ++ch) {
            const uint64_t cc = nc * 8 + ch;
            if (cc == c) continue;
            const int s = find_box(code, ls, depth, L, cc);
            if (s < 0) continue;
            int cx, cy, cz;
            morton_decode(cc, cx, cy, cz);
            if (max(abs(cx - bx), max(abs(cy - by), abs(cz - bz))) > 1) emit(0, s);
        }
        if (L > 1) {
            const int s = find_box(code, ls, depth, L - 1, nc);
            if (s >= 0 && is_leaf[s]) {
                const bool adj = adj_coarse(bx, (int)nx) && adj_coarse(by, (int)ny) && adj_coarse(bz, (int)nz);
                if (adj && leaf) emit(1, s);
                if (!adj) emit(3, s);
            }
        }
    }
    if (leaf) {
        for (int d = 0; d < 26; ++d) {
            int dx, dy, dz;
            neighbour_offset(d, dx, dy, dz);
            const int64_t mx = bx + dx, my = by + dy, mz = bz + dz;
            if (mx < 0 || my < 0 || mz < 0 || mx >= size || my >= size || mz >= size) continue;
            const uint64_t mc = morton_encode(mx, my, mz);
            const int s = find_box(code, ls, depth, L, mc);
            if (s >= 0 && is_leaf[s]) emit(1, s);
            if (L + 1 > depth) continue;
            for (int ch = 0; ch < 8; ++ch) {
                const uint64_t fc = mc * 8 + ch;
                const int f = find_box(code, ls, depth, L + 1, fc);
                if (f < 0) continue;
                int fx, fy, fz;
                morton_decode(fc, fx, fy, fz);
                const bool adj = adj_fine(bx, fx) && adj_fine(by, fy) && adj_fine(bz, fz);
                if (adj && is_leaf[f]) emit(1, f);
                if (!adj) emit(2, f);
            }
        }
    }
    if (!fill)
        for (int w = 0; w < 4; ++w) cnt[(int64_t)w * nb + b] = count[w];
}

struct DLevel {
    DBuf<uint64_t> code;
    DBuf<int> count;
    DBuf<uint8_t> leaf;
    int n = 0;
};

// boxes of every level for a fixed finest level md (the device form of build_fixed)
void build_levels(const double* d_xyz, int64_t n, int md, int n_max, const double* pmin, double extent, Temp& tmp,
                  std::vector<DLevel>& lv, DBuf<uint64_t>& scodes, DBuf<int>& sidx) {
    DBuf<uint64_t> codes(n);
    DBuf<int> idx(n);
    scodes.resize(n);
    sidx.resize(n);
    const double grid = std::ldexp(1.0, md);
    hipLaunchKernelGGL(codes_kernel, dim3(blocks(n)), dim3(256), 0, 0, d_xyz, n, pmin[0], pmin[1], pmin[2], extent,
                       grid, (1ull << md) - 1, codes.p, idx.p);
    tchk(hipGetLastError(), "codes kernel");
    size_t bytes = 0;
    const unsigned un = (unsigned)n, mbits = (unsigned)(3 * md);
    tchk(rocprim::radix_sort_pairs(nullptr, bytes, codes.p, scodes.p, idx.p, sidx.p, un, 0u, mbits), "sort size");
    tchk(rocprim::radix_sort_pairs(tmp.get(bytes), bytes, codes.p, scodes.p, idx.p, sidx.p, un, 0u, mbits), "sort");
    lv.clear();
    lv.resize(md + 2);
    DBuf<int> nruns(1);
    DLevel& fin = lv[md];
    fin.code.resize(n);
    fin.count.resize(n);
    tchk(rocprim::run_length_encode(nullptr, bytes, scodes.p, un, fin.code.p, fin.count.p, nruns.p), "rle size");
    tchk(rocprim::run_length_encode(tmp.get(bytes), bytes, scodes.p, un, fin.code.p, fin.count.p, nruns.p), "rle");
    fin.n = read1(nruns.p);
    fin.leaf.resize(fin.n);
    hipLaunchKernelGGL(fill_u8_kernel, dim3(blocks(fin.n)), dim3(256), 0, 0, fin.leaf.p, fin.n, (uint8_t)1);
    for (int L = md; L >= 2; --L) {
        DLevel& cur = lv[L];
        DBuf<uint64_t> keys(cur.n);
        hipLaunchKernelGGL(shift_kernel, dim3(blocks(cur.n)), dim3(256), 0, 0, cur.code.p, cur.n, keys.p);
        DLevel par;
        par.code.resize(cur.n);
        par.count.resize(cur.n);
        const size_t cn = (size_t)cur.n;
        tchk(rocprim::reduce_by_key(nullptr, bytes, keys.p, cur.count.p, cn, par.code.p, par.count.p, nruns.p,
                                    rocprim::plus<int>()),
             "reduce size");
        tchk(rocprim::reduce_by_key(tmp.get(bytes), bytes, keys.p, cur.count.p, cn, par.code.p, par.count.p, nruns.p,
                                    rocprim::plus<int>()),
             "reduce");
        par.n = read1(nruns.p);
        const uint64_t last_code = read1(cur.code.p + cur.n - 1);
        DBuf<uint8_t> merge(par.n);
        const DLevel& fine = lv[L + 1];
        hipLaunchKernelGGL(merge_kernel, dim3(blocks(par.n)), dim3(256), 0, 0, par.code.p, par.count.p, par.n,
                           cur.code.p, cur.leaf.p, cur.n, fine.code.p, fine.leaf.p, fine.n, last_code, L - 1,
                           L != md ? 1 : 0, n_max, merge.p);
        tchk(hipGetLastError(), "merge kernel");
        DBuf<uint8_t> keep(cur.n);
        hipLaunchKernelGGL(keep_kernel, dim3(blocks(cur.n)), dim3(256), 0, 0, cur.code.p, cur.n, par.code.p, par.n,
                           merge.p, keep.p);
        DLevel kept;
        kept.code.resize(cur.n);
        kept.count.resize(cur.n);
        kept.leaf.resize(cur.n);
        tchk(rocprim::select(nullptr, bytes, cur.code.p, keep.p, kept.code.p, nruns.p, cn), "select size");
        tchk(rocprim::select(tmp.get(bytes), bytes, cur.code.p, keep.p, kept.code.p, nruns.p, cn), "select");
        tchk(rocprim::select(nullptr, bytes, cur.count.p, keep.p, kept.count.p, nruns.p, cn), "select size");
        tchk(rocprim::select(tmp.get(bytes), bytes, cur.count.p, keep.p, kept.count.p, nruns.p, cn), "select");
        tchk(rocprim::select(nullptr, bytes, cur.leaf.p, keep.p, kept.leaf.p, nruns.p, cn), "select size");
        tchk(rocprim::select(tmp.get(bytes), bytes, cur.leaf.p, keep.p, kept.leaf.p, nruns.p, cn), "select");
        kept.n = read1(nruns.p);
        cur = std::move(kept);
        par.leaf = std::move(merge);
        lv[L - 1] = std::move(par);
    }
}

// sort the (target << 32 | source) keys of one list and split them
void finish_list(DBuf<uint64_t>& keys, int64_t m, int nb, Temp& tmp, std::vector<int>& tgt, std::vector<int>& src) {
    tgt.clear();
    src.clear();
    if (m == 0) return;
    DBuf<uint64_t> sorted(m);
    size_t bytes = 0;
    const int end_bit = 32 + bits_for((uint64_t)nb);
    const unsigned um = (unsigned)m, ebits = (unsigned)end_bit;
    tchk(rocprim::radix_sort_keys(nullptr, bytes, keys.p, sorted.p, um, 0u, ebits), "sort size");
    tchk(rocprim::radix_sort_keys(tmp.get(bytes), bytes, keys.p, sorted.p, um, 0u, ebits), "sort");
    const std::vector<uint64_t> h = sorted.host(m);
    tgt.resize(m);
    src.resize(m);
    for (int64_t i = 0; i < m; ++i) {
        tgt[i] = (int)(h[i] >> 32);
        src[i] = (int)(uint32_t)h[i];
    }
}

}  // namespace

void build_tree_device(const double* xyz, int64_t n, int n_max, int max_depth, Tree& t, Lists& ls) {
    if (n <= 0) fail(-2, "positions must be a nonempty finite (N,3) array");
    if (n >= (int64_t)1 << 31) fail(-2, "too many points for the GPU tree build");
    t = Tree();
    t.n_points = n;
    double pmax[3];
    for (int a = 0; a < 3; ++a) {
        t.pmin[a] = xyz[a];
        pmax[a] = xyz[a];
    }
    for (int64_t i = 1; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            t.pmin[a] = std::min(t.pmin[a], xyz[3 * i + a]);
            pmax[a] = std::max(pmax[a], xyz[3 * i + a]);
        }
    t.extent = std::max(pmax[0] - t.pmin[0], std::max(pmax[1] - t.pmin[1], pmax[2] - t.pmin[2]));
    if (!(t.extent > 0)) t.extent = 1.0;

    DBuf<double> d_xyz(3 * n);
    tchk(hipMemcpy(d_xyz.p, xyz, 3 * n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    Temp tmp;
    const bool given = max_depth > 0;
    int md = given ? max_depth
                   : std::max(2, (int)std::ceil(std::log(std::max((double)n / std::max(n_max, 1), 1.0)) /
                                                std::log(8.0)));
    const int depth_cap = 10;
    std::vector<DLevel> lv;
    DBuf<uint64_t> scodes;
    DBuf<int> sidx;
    std::vector<std::vector<uint64_t>> hcode;
    std::vector<std::vector<int>> hcount;
    std::vector<std::vector<uint8_t>> hleaf;
    int maxleaf = 0;
    while (true) {
        if (md > kMaxLevels - 2) fail(-2, "tree depth exceeds the supported maximum");
        build_levels(d_xyz.p, n, md, n_max, t.pmin, t.extent, tmp, lv, scodes, sidx);
        hcode.assign(md + 1, {});
        hcount.assign(md + 1, {});
        hleaf.assign(md + 1, {});
        maxleaf = 0;
        for (int L = 1; L <= md; ++L) {
            hcode[L] = lv[L].code.host(lv[L].n);
            hcount[L] = lv[L].count.host(lv[L].n);
            hleaf[L] = lv[L].leaf.host(lv[L].n);
            for (int i = 0; i < lv[L].n; ++i)
                if (hleaf[L][i]) maxleaf = std::max(maxleaf, hcount[L][i]);
        }
        if (given || maxleaf <= 2 * n_max || md >= depth_cap) break;
        ++md;
    }
    t.depth = md;
    t.maxleaf = maxleaf;
    t.level_start.assign(md + 2, 0);
    for (int L = 1; L <= md; ++L) {
        t.level_start[L] = (int)t.level.size();
        for (size_t i = 0; i < hcode[L].size(); ++i) {
            t.level.push_back(L);
            t.code.push_back(hcode[L][i]);
            t.count.push_back(hcount[L][i]);
            t.is_leaf.push_back(hleaf[L][i]);
        }
    }
    t.level_start[md + 1] = (int)t.level.size();
    t.level_start[0] = 0;
    const int nb = t.n_boxes();

    DBuf<uint64_t> d_code;
    DBuf<int> d_level, d_ls, d_pstart(nb), d_parent(nb);
    DBuf<uint8_t> d_leaf;
    d_code.upload(t.code);
    d_level.upload(t.level);
    d_ls.upload(t.level_start);
    d_leaf.upload(t.is_leaf);
    hipLaunchKernelGGL(box_kernel, dim3(blocks(nb)), dim3(256), 0, 0, d_code.p, d_level.p, d_ls.p, md, nb, scodes.p, n,
                       d_pstart.p, d_parent.p);
    tchk(hipGetLastError(), "box kernel");
    t.pstart = d_pstart.host(nb);
    t.parent = d_parent.host(nb);
    const std::vector<int> perm = sidx.host(n);
    t.perm.assign(perm.begin(), perm.end());

    // lists: count, scan, fill, sort
    DBuf<int> cnt((size_t)4 * nb);
    hipLaunchKernelGGL(lists_kernel, dim3(blocks(nb)), dim3(256), 0, 0, d_code.p, d_level.p, d_leaf.p, d_ls.p, md, nb,
                       cnt.p, (const int64_t*)nullptr, (uint64_t*)nullptr, (uint64_t*)nullptr, (uint64_t*)nullptr,
                       (uint64_t*)nullptr);
    tchk(hipGetLastError(), "lists count kernel");
    const std::vector<int> hc = cnt.host((size_t)4 * nb);
    std::vector<int64_t> hoff((size_t)4 * nb);
    int64_t tot[4];
    for (int w = 0; w < 4; ++w) {
        int64_t s = 0;
        for (int b = 0; b < nb; ++b) {
            hoff[(size_t)w * nb + b] = s;
            s += hc[(size_t)w * nb + b];
        }
        tot[w] = s;
    }
    DBuf<int64_t> d_off;
    d_off.upload(hoff);
    DBuf<uint64_t> out[4];
    for (int w = 0; w < 4; ++w) out[w].resize(std::max<int64_t>(tot[w], 1));
    hipLaunchKernelGGL(lists_kernel, dim3(blocks(nb)), dim3(256), 0, 0, d_code.p, d_level.p, d_leaf.p, d_ls.p, md, nb,
                       cnt.p, d_off.p, out[0].p, out[1].p, out[2].p, out[3].p);
    tchk(hipGetLastError(), "lists fill kernel");
    ls = Lists();
    finish_list(out[0], tot[0], nb, tmp, ls.v_tgt, ls.v_src);
    finish_list(out[1], tot[1], nb, tmp, ls.u_tgt, ls.u_src);
    finish_list(out[2], tot[2], nb, tmp, ls.w_tgt, ls.w_src);
    finish_list(out[3], tot[3], nb, tmp, ls.x_tgt, ls.x_src);
    tchk(hipDeviceSynchronize(), "tree build");
}

}  // namespace hfmm
