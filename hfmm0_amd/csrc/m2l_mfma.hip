// M2L on matrix cores: all pairs of a block of target boxes that share a translation
// operator are contracted together.  The same kernel runs M2M (targets: parents, sources:
// their children, source degree = child level's) and L2L (targets: children of one octant,
// source: their parents) with source and target degrees that differ.
//
// A workgroup (4 waves) owns TB target boxes of one level that have the same position
// (octant) inside their parents, so their V lists use the same set of displacements.
// Displacements with the same polar angle theta and length |d| (they differ only in the
// azimuth phi) share the rotation E_n(theta) and the coaxial operator T(|d|), so they are
// processed together as one group of up to G directions: G * TB pairs, 2 G TB columns.
// For each group the workgroup
//   0. gathers the G * TB source multipoles into LDS, each multiplied by e^{i m phi} of its
//      direction (rows = coefficient n*n+n+m, columns = (direction, pair, re/im));
//   1. rotates each degree block in place: X_n <- E_n^T X_n   (real x complex);
//   2. applies the coaxial operator per order m in place: X^m <- T^{|m|} X^m, as
//      Re(T) * X + Im(T) * (X with re/im columns swapped and signed);
//   3. rotates back, W_n = E_n X_n, multiplies by e^{-i m phi} of each column's direction
//      (the partner lane holds the other half of the complex value) and accumulates per
//      target in registers.
// Every product is a 16x16x4 MFMA tile.  The A operands come from operator tables that
// the host stores in fragment order, zero-padded to whole 16x16 tiles, so a lane loads
// the A values of four k-steps with one 16-byte (FP32) or two 16-byte (FP64) loads and
// no bounds test; each A fragment is applied to every 16-column tile of the group.  B
// rows read past a block hold finite values (zero, or results of earlier groups) whose
// products with the zero padding of A vanish; source degrees above the evaluated one read
// a zero row; results of rows outside a block are stored to a scratch row.  The stage-3
// accumulators stay in registers across all groups and are added to the FP64 local
// expansions once at the end.  Work units (degrees, orders, row tiles) are dealt to the
// four waves by a host-built list that balances their MFMA counts.
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"

namespace hfmm {

// LDS: the expansion block (with its scratch row), then the phase table: 2 lane parities
// x G directions x m2l_phase_len(p) complex entries
size_t m2l_block_lds_bytes(int p, int tb, int g, int precision_bytes) {
    return ((size_t)m2l_lds_x_elems(p, 2 * tb * g) + 4 * (size_t)g * m2l_phase_len(p)) * precision_bytes;
}
// p: the larger of the source and target table degrees

namespace {

template <typename T>
struct Mma;
template <>
struct Mma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    typedef float2 c2;
    __device__ static inline v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // 4 rows x 64 columns, K = 1: the 4x4x1 MFMA over 16 blocks with the A operand of block 0
    // (lanes 0..3, row = lane) broadcast to all blocks; b: lane l holds column l, and output
    // element r of lane l is row r of column l.  Same FP32 rate as 16x16x4, 8 passes.
    __device__ static inline v4 mma4(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 4, 0, 0);
    }
    // output element r of lane group lk (= lane / 16) is row 4 lk + r of the 16x16 tile
    __host__ __device__ static constexpr int row(int lk, int r) { return 4 * lk + r; }
    __device__ static inline v4 load4(const float* p) { return *(const v4*)p; }
    // value of the neighbouring lane (lane ^ 1): quad permutation (1, 0, 3, 2)
    __device__ static inline float swap_pair(float v) {
        return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
    }
};
template <>
struct Mma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    typedef double2 c2;
    __device__ static inline v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // output element r of lane group lk (= lane / 16) is row lk + 4 r of the 16x16 tile
    __host__ __device__ static constexpr int row(int lk, int r) { return lk + 4 * r; }
    __device__ static inline v4 load4(const double* p) {
        const double2 a = *(const double2*)p, b = *(const double2*)(p + 2);
        return v4{a.x, a.y, b.x, b.y};
    }
    __device__ static inline double swap_pair(double v) {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_mov_dpp((int)b, 0xB1, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, false);
        return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
    }
};

// index q of the pair (n, m >= 0), q = n (n+1) / 2 + m -> (n, m).  With x = 8 q + 1.5,
// (2n+1)^2 + 1/2 <= x <= (2n+3)^2 - 15/2, so sqrt(x) lies at least 1/(4 (2n+3)) inside
// (2n+1, 2n+3); for q < 2^16 the hardware square root (1 ulp) gives
// n = floor((sqrt(x) - 1) / 2).
__device__ inline void decode_q(int q, int& n, int& m) {
    n = (int)((__builtin_amdgcn_sqrtf(8.0f * (float)q + 1.5f) - 1.0f) * 0.5f);
    m = q - ((n * (n + 1)) >> 1);
}

// Work item: groups [g0, g1) of one target block.  When a block is split over several
// items (coarse levels, to fill the GPU) the results are added atomically.  KT is the
// number of 16-row tiles of the largest (u or v) block, m2l_tiles(p_tab + 1) (1 up to
// p = 15, 2 up to 23); it bounds the A fragments a unit keeps in registers.
// waves per SIMD the compiler fits the registers to: FP32 with 16 x 2 pairs per group and
// one tile per block (the finest levels) is held to 4 (128 VGPRs; 4.31 ms at p = 11,
// N = 1M, k = 40, against 4.69 at the 3 waves its registers would otherwise take).  It fits
// without spilling because the lane-derived offsets are formed again in every group and the
// gather's source offsets are scalar.  The other instances keep the compiler's choice
// (FP64 at 4 spills heavily and runs 35% slower).
template <typename T, int TB, int G, int KT>
constexpr int m2l_block_min_waves() {
    return sizeof(T) == 4 && TB == 16 && G == 2 && KT == 1 ? 4 : 1;
}

// One work item (block, group range) of the translation described by a.
template <typename T, int TB, int G, int KT, int MAXU>
__device__ __forceinline__ void m2l_block_body(const M2LBlockArgs& a, const int4 item) {
    using V4 = typename Mma<T>::v4;
    using C2 = typename Mma<T>::c2;
    constexpr int kMaxTiles = KT;           // row tiles of a rotation block or an order block
    constexpr int NCOLD = 2 * TB;           // columns of one direction: (pair, re/im)
    constexpr int NCTD = NCOLD / 16;        // 16-wide column tiles of one direction
    constexpr int NCT = G * NCTD;           // column tiles of the group
    constexpr int LDX = G * NCOLD + 1;      // odd row stride in LDS
    // phase table: per lane parity and direction, entries for the orders m = 0 .. TL-1:
    // (cos m phi, +-sin m phi), the sine negated for odd (imaginary) columns; orders
    // m >= PH hold zero
    constexpr int TL = 16 * KT;
    // TB = 16: the real parts of a direction's 16 pairs fill one 16-column tile and the
    // imaginary parts the next (SPL), so a value's partner (re <-> im) sits in the same lane
    // of the other tile; otherwise (TB = 8) the columns of a pair are adjacent (re, im) and
    // the partner is the neighbouring lane
    constexpr bool SPL = TB == 16;
    constexpr int RSTEP = Mma<T>::row(0, 1) - Mma<T>::row(0, 0);   // rows between output registers
    // FP32 with one tile per block (p_tab <= 15, the finest levels): every unit of every
    // stage uses one fragment set, and the next unit's set is loaded while the current one
    // is multiplied (the L2 latency of the operator tables is otherwise exposed per unit)
    constexpr bool PF = sizeof(T) == 4 && KT == 1;
    // with PF, blocks of at most kM2LQuadMaxRows rows run on the 4x4x1 MFMA in 4-row tiles
    // (lane = column; a 32-column block reads its columns twice and writes from the first
    // 32 lanes); the code takes up to two tiles
    static_assert(kM2LQuadMaxRows <= 8, "4x4x1 path: at most two 4-row tiles");
    constexpr int QT = (kM2LQuadMaxRows + 3) / 4, QR = 4 * QT;   // tiles and rows of the path
    constexpr int NCOLS = G * NCOLD;
    constexpr bool Q4 = PF && (NCOLS == 64 || NCOLS == 32);
    extern __shared__ char smem_raw[];
    // source degree p (multipoles or parent locals) and target degree pt; equal for M2L
    const int p = a.p, P1 = p + 1, nc = P1 * P1, nq = P1 * (P1 + 1) / 2;
    const int pt = a.pt, PT1 = pt + 1, nct = PT1 * PT1;
    const int ptab = max(a.p_tab, a.pt_tab);           // LDS rows are sized for this degree
    const int PH = max(P1, PT1);                        // phases per direction
    const int spare = (ptab + 1) * (ptab + 1);          // scratch row for out-of-block stores
    T* X = (T*)smem_raw;
    C2* Ps = (C2*)(X + m2l_lds_x_elems(ptab, G * NCOLD));   // [parity][G][TL]
    // the wave index is uniform across a wavefront: keep it (and all unit bookkeeping
    // derived from it) in scalar registers so that the unit branches are scalar
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int li = lane & 15, lk = lane >> 4;
    // this lane's phase entries of stage 3: add 16 mt for a unit's row tile mt and
    // g TL + r RSTEP for direction g and output register r
    const C2* ps_lane = Ps + (SPL ? 0 : (li & 1)) * G * TL + Mma<T>::row(lk, 0);
    if (item.y >= item.z) return;   // padding item
    const int* tg = a.blk_targets + (size_t)item.x * TB;
    const C2* Mexp = (const C2*)a.M;

    // stage-3 accumulators of the wave's units in the (u, v) basis, summed over all groups
    V4 accu[MAXU][NCTD], accv[MAXU][NCTD];
#pragma unroll
    for (int u = 0; u < MAXU; ++u)
#pragma unroll
        for (int cd = 0; cd < NCTD; ++cd) {
            accu[u][cd] = V4{0, 0, 0, 0};
            accv[u][cd] = V4{0, 0, 0, 0};
        }
    // zero the whole block once: rows that the padded tiles read beyond the gathered
    // degrees are then finite (stale results of earlier groups) or zero, and their
    // products with the zero padding of A vanish
    const int zero_row = spare + 1;   // the rows after the scratch row stay zero
    for (int i = threadIdx.x; i < m2l_lds_x_rows(ptab) * LDX; i += 256) X[i] = (T)0;
    // with PF: the wave's units of each stage in its lanes (lane i: unit wave + 4 i), as
    // fragment offset << 9 | code, 255 for a unit to skip; read with v_readlane, so stepping
    // through the units needs no scalar memory load (whose wait would also drain the LDS
    // queue)
    const int nu1 = a.n_units1 >> 2, nu2 = a.n_units2 >> 2;   // the lists are padded to rows of 4
    int vt1 = 255, vt2 = 255, vt3 = 255;
    if constexpr (PF) {
        // blocks of at most kM2LQuadMaxRows rows (forward rotation; square, so the size is
        // fixed by the degree) and coaxial blocks of at most that size at the table degrees
        // take the 4x4x1 fragments (bit 8 of a coaxial unit's code marks them)
        if (lane < nu1) {
            const int c = a.units1[wave + 4 * lane];
            if ((c & 255) <= p) {
                const int b = 2 * (c & 255) + (c >> 8), oq = Q4 ? a.rot_off1q[b] : -1;
                vt1 = ((oq >= 0 ? oq : a.rot_off1[b]) << 9) | c;
            }
        }
        if (lane < nu2) {
            const int c = a.units2[wave + 4 * lane];
            if (c <= p && c <= pt) {
                const int oq = Q4 ? a.coax_offq[c] : -1;
                vt2 = oq >= 0 ? (oq << 9) | 256 | c : (a.coax_off[c] << 9) | c;
            }
        }
        if (lane < MAXU && wave + 4 * lane < a.n_units3) {
            const int c = a.units3[wave + 4 * lane];
            if (c & 256) {   // packed degrees (n1, n2): active while n1 is
                if (((c >> 4) & 15) <= pt) vt3 = (a.rot_off3p[8 * ((c >> 4) & 15) + (c & 15)] << 9) | c;
            } else if (c <= pt) {
                vt3 = (a.rot_off3[2 * c] << 9) | c;   // KT = 1: row tile 0, code = n
            }
        }
    }
    auto skip = [](int t) { return (t & 511) == 255; };
    __syncthreads();

    for (int e = item.y; e < item.z; ++e) {
        // FP32 one-tile instances (4 waves per SIMD, 128 registers) form the lane-derived
        // offsets again in every group from an opaque copy of the lane index: kept live across
        // the loop they would be spilled and reloaded from scratch in every group.  The other
        // instances keep them (the FP64 ones ran 14% slower with the recomputation).
        int zero_o = 0;
        if constexpr (PF) asm volatile("" : "+v"(zero_o));
        const int lane_o =
            PF ? __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, (unsigned)zero_o)) : lane;
        const int li = lane_o & 15, lk = lane_o >> 4;
        const C2* ps_lane = Ps + (SPL ? 0 : (li & 1)) * G * TL + Mma<T>::row(lk, 0);
        const unsigned short* gd = a.grp_dir + (size_t)e * G;
        const DispInfo di = a.dinfo[gd[0]];   // shared by the directions of the group
        const T* R1 = (const T*)a.rot1 + (size_t)di.theta * a.rot_stride1;
        const T* R3 = (const T*)a.rot3 + (size_t)di.theta * a.rot_stride3;
        const T* Tz = (const T*)a.coax + (size_t)di.dist * a.coax_stride;
        const int* soff = a.grp_off + (size_t)e * G * TB;   // -1: no source (or no direction)

        // stage 0: phase table; gather, times e^{i m phi}, into the (u, v) basis
        for (int i = threadIdx.x; i < 2 * G * TL; i += 256) {
            const int par = i / (G * TL), g = (i / TL) % G, m = i % TL;
            const unsigned short d = gd[g];
            C2 v;
            v.x = 0;
            v.y = 0;
            if (m < PH) {
                v.x = 1;
                if (d != 0xFFFF) v = ((const C2*)a.phase)[(size_t)d * a.phase_stride + m];
                if (par) v.y = -v.y;
            }
            Ps[i] = v;
        }
        // every wave gathers for one direction (uniform source offsets, scalar loads),
        // lanes over the pairs (n, m >= 0): x_m e^{i m phi} and x_{-m} e^{-i m phi} of each
        // source give u_m (row of x_m) and v_m (row of x_{-m})
        {
            constexpr int WPD = 4 / G;   // waves per direction
            const int g = wave / WPD;
            const unsigned short d = gd[g];
            const int* so = soff + g * TB;
            for (int q = (wave % WPD) * 64 + lane_o; q < nq; q += 64 * WPD) {
                int n, m;
                decode_q(q, n, m);
                const int cp = n * n + n + m, cm = n * n + n - m;
                C2 pm;
                pm.x = 1;
                pm.y = 0;
                if (d != 0xFFFF) pm = ((const C2*)a.phase)[(size_t)d * a.phase_stride + m];
                const T sg = (m & 1) ? (T)-1 : (T)1;
                T* xu = X + cp * LDX + g * NCOLD;
                T* xv = X + cm * LDX + g * NCOLD;
                // sources in batches of JB (all loads of a batch in flight together; half of
                // the block when the split layout's separate re / im stores keep more live)
                constexpr int JB = sizeof(T) == 8 || SPL ? TB / 2 : TB;
#pragma unroll
                for (int j0 = 0; j0 < TB; j0 += JB) {
                    C2 rp[JB], rm[JB];
#pragma unroll
                    for (int j = 0; j < JB; ++j) {
                        // wave-uniform: a scalar base for the lanes' coefficient offsets
                        const int o = __builtin_amdgcn_readfirstlane(so[j0 + j]);
                        rp[j].x = 0;
                        rp[j].y = 0;
                        rm[j].x = 0;
                        rm[j].y = 0;
                        if (o >= 0) {
                            rp[j] = Mexp[(int64_t)o + cp];
                            if (m > 0) rm[j] = Mexp[(int64_t)o + cm];
                        }
                    }
#pragma unroll
                    for (int j = 0; j < JB; ++j) {
                        // a = x_m e^{i m phi}, b = (-1)^m x_{-m} e^{-i m phi}
                        const T ar = rp[j].x * pm.x - rp[j].y * pm.y, ai = rp[j].x * pm.y + rp[j].y * pm.x;
                        const T br = sg * (rm[j].x * pm.x + rm[j].y * pm.y);
                        const T bi = sg * (rm[j].y * pm.x - rm[j].x * pm.y);
                        const int cre = SPL ? j0 + j : 2 * (j0 + j), cim = SPL ? TB + j0 + j : 2 * (j0 + j) + 1;
                        xu[cre] = ar + br;
                        xu[cim] = ai + bi;
                        if (m > 0) {
                            xv[cre] = ar - br;
                            xv[cim] = ai - bi;
                        }
                    }
                }
            }
            // target degrees above the source degree start from zero (orders |m| > p
            // receive nothing in stage 2)
            for (int c = nc + (wave % WPD) * 64 + lane_o; c < nct; c += 64 * WPD) {
                T* xr = X + c * LDX + g * NCOLD;
#pragma unroll
                for (int j = 0; j < NCOLD; ++j) xr[j] = (T)0;
            }
        }

        // stage 1: forward rotation in place, per degree n and block (u: rows n^2+n ..
        // n^2+2n, v: rows n^2 .. n^2+n-1), X_b <- W_b X_b.  A unit is one block with all
        // its row and column tiles: the A fragments of a k group are applied to every
        // column tile; the wave writes back only after all products of the unit are formed,
        // and no other wave touches the rows of the block.
        if constexpr (PF) {
            auto nx = [&](int i) {
                while (i < nu1 && skip(__builtin_amdgcn_readlane(vt1, i))) ++i;
                return i;
            };
            int i1 = nx(0);
            V4 an = V4{0, 0, 0, 0};
            if (i1 < nu1) an = Mma<T>::load4(R1 + (__builtin_amdgcn_readlane(vt1, i1) >> 9) + lane_o * 4);
            __syncthreads();
            while (i1 < nu1) {
                const int code1 = __builtin_amdgcn_readlane(vt1, i1);
                const int code = code1 & 511;
                const int n = code & 255, h = code >> 8;
                const int len = h ? n : n + 1, base = h ? n * n : n * n + n, nks = (len + 3) >> 2;
                const V4 av = an;
                i1 = nx(i1 + 1);
                if (i1 < nu1) an = Mma<T>::load4(R1 + (__builtin_amdgcn_readlane(vt1, i1) >> 9) + lane_o * 4);
                if (Q4 && len <= kM2LQuadMaxRows) {
                    // 4-row tiles: lane = column, k = 0..len-1 over the block's rows; the
                    // fragment of tile t and k group g at (t nt4 + g) x 256 (the first is av)
                    const int cl = lane_o & (NCOLS - 1), nt4 = (len + 3) >> 2;
                    const T* A4 = R1 + (code1 >> 9) + lane_o * 4;
                    T b4[QR];
#pragma unroll
                    for (int k = 0; k < QR; ++k) b4[k] = X[(base + k) * LDX + cl];   // past len: A is 0
#pragma unroll
                    for (int t4 = 0; t4 < QT; ++t4) {
                        if (t4 >= nt4) break;
                        V4 d4 = V4{0, 0, 0, 0};
#pragma unroll
                        for (int g4 = 0; g4 < QT; ++g4) {
                            if (g4 >= nt4) break;
                            const V4 a4 = t4 + g4 == 0 ? av : Mma<T>::load4(A4 + (t4 * nt4 + g4) * 256);
#pragma unroll
                            for (int j = 0; j < 4; ++j)
                                if (4 * g4 + j < len) d4 = Mma<T>::mma4(a4[j], b4[4 * g4 + j], d4);
                        }
                        if (lane_o < NCOLS)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                if (4 * t4 + r < len) X[(base + 4 * t4 + r) * LDX + cl] = d4[r];
                    }
                    continue;
                }
                V4 c4[NCT];
                T bv[4][NCT];
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct) c4[ct] = V4{0, 0, 0, 0};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int row = base + 4 * j + lk;   // past len: A is 0, X is finite
#pragma unroll
                    for (int ct = 0; ct < NCT; ++ct) bv[j][ct] = X[row * LDX + ct * 16 + li];
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (j < nks)
#pragma unroll
                        for (int ct = 0; ct < NCT; ++ct) c4[ct] = Mma<T>::mma(av[j], bv[j][ct], c4[ct]);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = Mma<T>::row(lk, r);
                    const int dst = __umul24(row < len ? base + row : spare, LDX);
#pragma unroll
                    for (int ct = 0; ct < NCT; ++ct) X[dst + ct * 16 + li] = c4[ct][r];
                }
            }
        } else {
        __syncthreads();
        for (int u = wave; u < a.n_units1; u += 4) {
            const int code = a.units1[u];
            const int n = code & 255, h = code >> 8;
            if (n > p) continue;
            const int len = h ? n : n + 1, base = h ? n * n : n * n + n;
            const int nt = m2l_tiles(len), nks = (len + 3) >> 2;
            const T* A = R1 + a.rot_off1[2 * n + h];
            V4 c4[kMaxTiles][NCT];
#pragma unroll
            for (int mt = 0; mt < kMaxTiles; ++mt)
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct) c4[mt][ct] = V4{0, 0, 0, 0};
            for (int kg = 0; kg < nt; ++kg) {
                V4 av[kMaxTiles];
#pragma unroll
                for (int mt = 0; mt < kMaxTiles; ++mt)
                    if (mt < nt) av[mt] = Mma<T>::load4(A + ((size_t)(mt * nt + kg) * 64 + lane_o) * 4);
                T bv[4][NCT];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int row = base + 16 * kg + 4 * j + lk;   // past len: A is 0, X is finite
#pragma unroll
                    for (int ct = 0; ct < NCT; ++ct) bv[j][ct] = X[row * LDX + ct * 16 + li];
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (4 * kg + j < nks)
#pragma unroll
                        for (int mt = 0; mt < kMaxTiles; ++mt)
                            if (mt < nt)
#pragma unroll
                                for (int ct = 0; ct < NCT; ++ct)
                                    c4[mt][ct] = Mma<T>::mma(av[mt][j], bv[j][ct], c4[mt][ct]);
            }
#pragma unroll
            for (int mt = 0; mt < kMaxTiles; ++mt)
                if (mt < nt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = mt * 16 + Mma<T>::row(lk, r);
                        const int dst = __umul24(row < len ? base + row : spare, LDX);
#pragma unroll
                        for (int ct = 0; ct < NCT; ++ct) X[dst + ct * 16 + li] = c4[mt][ct][r];
                    }
        }

        }
        // stage 2: X^m <- T^{|m|} X^m in place on the u_m and the v_m (the operator is the
        // same for the orders +m and -m, hence for their sum and difference).  A unit is
        // one m with both blocks: its A fragments of a k group are applied to both and to
        // every column tile.  The table blocks are built for p_tab; source degrees beyond
        // the evaluated p read a zero row.
        {
            const T sgn1 = (li & 1) ? (T)1 : (T)-1;   // Im(T) x X: (-Im X, Re X) per column pair
            if constexpr (PF) {
                auto nx = [&](int i) {
                    while (i < nu2 && skip(__builtin_amdgcn_readlane(vt2, i))) ++i;
                    return i;
                };
                V4 anr = V4{0, 0, 0, 0}, ani = V4{0, 0, 0, 0};
                auto pre = [&](int i) {   // fragment (row tile 0, k group 0) of the order's block
                    const T* At = Tz + (__builtin_amdgcn_readlane(vt2, i) >> 9) + lane_o * 8;
                    anr = Mma<T>::load4(At);
                    ani = Mma<T>::load4(At + 4);
                };
                int i2 = nx(0);
                if (i2 < nu2) pre(i2);
                __syncthreads();
                while (i2 < nu2) {
                    const int code2 = __builtin_amdgcn_readlane(vt2, i2);
                    const int am = code2 & 255;
                    const int ns = am > 0 ? 2 : 1;
                    const int len = P1 - am, lent = PT1 - am, nks = (len + 3) >> 2;
                    const V4 are = anr, aim = ani;
                    i2 = nx(i2 + 1);
                    if (i2 < nu2) pre(i2);
                    if (Q4 && (code2 & 256)) {
                        // block of at most kM2LQuadMaxRows squared (table degrees) in 4-row tiles: lane = column;
                        // Re T on the column, Im T on its partner (re <-> im), signed for the
                        // real columns.  Fragment of tile t and k group g at (t nk4 + g) x 512
                        // (the first is are, aim).
                        constexpr int PART = SPL ? 16 : 1;
                        const int cl = lane_o & (NCOLS - 1);
                        const T sgi = (cl & PART) ? (T)1 : (T)-1;
                        const int nt4 = (a.pt_tab + 4 - am) >> 2, nk4 = (a.p_tab + 4 - am) >> 2;
                        const T* A4 = Tz + (code2 >> 9) + lane_o * 8;
#pragma unroll
                        for (int sg = 0; sg < 2; ++sg) {
                            if (sg < ns) {
                                T b1[QR], b2[QR];
#pragma unroll
                                for (int k = 0; k < QR; ++k) {
                                    const int nn = am + k;
                                    const int row = k < len ? nn * (nn + 1) + (sg ? -am : am) : zero_row;
                                    b1[k] = X[row * LDX + cl];
                                    b2[k] = X[row * LDX + (cl ^ PART)];
                                }
#pragma unroll
                                for (int t4 = 0; t4 < QT; ++t4) {
                                    if (t4 >= nt4 || 4 * t4 >= lent) break;
                                    V4 d1 = V4{0, 0, 0, 0}, d2 = V4{0, 0, 0, 0};
#pragma unroll
                                    for (int g4 = 0; g4 < QT; ++g4) {
                                        if (g4 >= nk4) break;
                                        V4 ar = are, ai = aim;
                                        if (t4 + g4 > 0) {
                                            ar = Mma<T>::load4(A4 + (t4 * nk4 + g4) * 512);
                                            ai = Mma<T>::load4(A4 + (t4 * nk4 + g4) * 512 + 4);
                                        }
#pragma unroll
                                        for (int j = 0; j < 4; ++j)
                                            if (4 * g4 + j < len) {
                                                d1 = Mma<T>::mma4(ar[j], b1[4 * g4 + j], d1);
                                                d2 = Mma<T>::mma4(ai[j], b2[4 * g4 + j], d2);
                                            }
                                    }
                                    if (lane_o < NCOLS)
#pragma unroll
                                        for (int r = 0; r < 4; ++r)
                                            if (4 * t4 + r < lent) {
                                                const int nn = am + 4 * t4 + r;
                                                X[(nn * (nn + 1) + (sg ? -am : am)) * LDX + cl] =
                                                    fma(sgi, d2[r], d1[r]);
                                            }
                                }
                            }
                        }
                        continue;
                    }
                    const int q0 = __umul24(am + lk, am + lk + 1);
                    int rowp[4], rowm[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int q = q0 + __umul24(8 * j, am + lk) + 16 * j * j + 4 * j;
                        const bool ok = 4 * j + lk < len;
                        rowp[j] = __umul24(ok ? q + am : zero_row, LDX);
                        rowm[j] = __umul24(ok ? q - am : zero_row, LDX);
                    }
                    const int nr0 = am + Mma<T>::row(lk, 0);
                    const int qr = __umul24(nr0, nr0 + 1);
                    // the order +m (sg 0), then -m: each result is stored before the other
                    // order's rows are read (the rows n^2+n+m and n^2+n-m are disjoint), so
                    // only one set of output tiles is live
#pragma unroll
                    for (int sg = 0; sg < 2; ++sg) {
                        if (sg < ns) {
                            V4 c4[NCT];
#pragma unroll
                            for (int ct = 0; ct < NCT; ++ct) c4[ct] = V4{0, 0, 0, 0};
                            if constexpr (SPL) {
                                // re tile: Re T x - Im T y, im tile: Re T y + Im T x, with the
                                // partner values in the same lane of the other tile
                                T b0[4][NCT];
#pragma unroll
                                for (int j = 0; j < 4; ++j) {
                                    const int cin = sg ? rowm[j] : rowp[j];
#pragma unroll
                                    for (int ct = 0; ct < NCT; ++ct) b0[j][ct] = X[cin + ct * 16 + li];
                                }
                                const V4 nim = -aim;
#pragma unroll
                                for (int j = 0; j < 4; ++j)
                                    if (j < nks)
#pragma unroll
                                        for (int ct = 0; ct < NCT; ++ct) {
                                            c4[ct] = Mma<T>::mma(are[j], b0[j][ct], c4[ct]);
                                            c4[ct] = Mma<T>::mma((ct & 1) ? aim[j] : nim[j], b0[j][ct ^ 1], c4[ct]);
                                        }
                            } else {
                            T b0[4][NCT], b1[4][NCT];
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                const int cin = sg ? rowm[j] : rowp[j];
#pragma unroll
                                for (int ct = 0; ct < NCT; ++ct) {
                                    b0[j][ct] = X[cin + ct * 16 + li];
                                    b1[j][ct] = sgn1 * X[cin + ct * 16 + (li ^ 1)];
                                }
                            }
#pragma unroll
                            for (int j = 0; j < 4; ++j)
                                if (j < nks)
#pragma unroll
                                    for (int ct = 0; ct < NCT; ++ct) {
                                        c4[ct] = Mma<T>::mma(are[j], b0[j][ct], c4[ct]);
                                        c4[ct] = Mma<T>::mma(aim[j], b1[j][ct], c4[ct]);
                                    }
                            }
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int sr = r * RSTEP;
                                const int q = qr + __umul24(2 * sr, nr0) + sr * sr + sr;
                                const bool ok = Mma<T>::row(lk, r) < lent;
                                const int dst = __umul24(ok ? (sg ? q - am : q + am) : spare, LDX);
#pragma unroll
                                for (int ct = 0; ct < NCT; ++ct) X[dst + ct * 16 + li] = c4[ct][r];
                            }
                        }
                    }
                }
            } else {
                __syncthreads();
                for (int u = wave; u < a.n_units2; u += 4) {
                    const int am = a.units2[u];
                    if (am > p || am > pt) continue;
                    const int ns = am > 0 ? 2 : 1;                  // u_m and v_m
                    const int len = P1 - am, lent = PT1 - am;       // source and target block sizes
                    const int ntab = m2l_tiles(a.p_tab + 1 - am);   // k groups of the table block
                    const int nt = m2l_tiles(lent), nkg = m2l_tiles(len), nks = (len + 3) >> 2;
                    const T* A = Tz + a.coax_off[am];
                    V4 c4[2][kMaxTiles][NCT];
    #pragma unroll
                    for (int sg = 0; sg < 2; ++sg)
    #pragma unroll
                        for (int t = 0; t < kMaxTiles; ++t)
    #pragma unroll
                            for (int ct = 0; ct < NCT; ++ct) c4[sg][t][ct] = V4{0, 0, 0, 0};
                    for (int kg = 0; kg < nkg; ++kg) {
                        V4 are[kMaxTiles], aim[kMaxTiles];
    #pragma unroll
                        for (int t = 0; t < kMaxTiles; ++t)
                            if (t < nt) {
                                const T* At = A + ((size_t)(t * ntab + kg) * 64 + lane_o) * 8;
                                are[t] = Mma<T>::load4(At);
                                aim[t] = Mma<T>::load4(At + 4);
                            }
                        // source rows of this k group: degree nn_j = am + 16 kg + 4 j + lk, row
                        // nn_j (nn_j + 1) +- am; nn_j (nn_j + 1) = q0 + 8 j nn_0 + 16 j^2 + 4 j
                        const int nn0 = am + 16 * kg + lk;
                        const int q0 = __umul24(nn0, nn0 + 1);
                        int rowp[4], rowm[4];
    #pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int q = q0 + __umul24(8 * j, nn0) + 16 * j * j + 4 * j;
                            const bool ok = 16 * kg + 4 * j + lk < len;
                            rowp[j] = __umul24(ok ? q + am : zero_row, LDX);
                            rowm[j] = __umul24(ok ? q - am : zero_row, LDX);
                        }
    #pragma unroll
                        for (int sg = 0; sg < 2; ++sg) {
                            if (sg < ns) {
                                T b0[4][NCT], b1[4][NCT];
    #pragma unroll
                                for (int j = 0; j < 4; ++j) {
                                    const int cin = sg ? rowm[j] : rowp[j];
    #pragma unroll
                                    for (int ct = 0; ct < NCT; ++ct) {
                                        b0[j][ct] = X[cin + ct * 16 + li];
                                        if constexpr (!SPL) b1[j][ct] = sgn1 * X[cin + ct * 16 + (li ^ 1)];
                                    }
                                }
                                if constexpr (SPL) {
                                    // partner values in the same lane of the other tile; the
                                    // re tiles take -Im T
    #pragma unroll
                                    for (int j = 0; j < 4; ++j)
    #pragma unroll
                                        for (int ct = 0; ct < NCT; ++ct) b1[j][ct] = b0[j][ct ^ 1];
                                }
    #pragma unroll
                                for (int j = 0; j < 4; ++j)
                                    if (4 * kg + j < nks)
    #pragma unroll
                                        for (int t = 0; t < kMaxTiles; ++t)
                                            if (t < nt)
    #pragma unroll
                                                for (int ct = 0; ct < NCT; ++ct) {
                                                    const T ai = SPL && !(ct & 1) ? -aim[t][j] : aim[t][j];
                                                    c4[sg][t][ct] = Mma<T>::mma(are[t][j], b0[j][ct], c4[sg][t][ct]);
                                                    c4[sg][t][ct] = Mma<T>::mma(ai, b1[j][ct], c4[sg][t][ct]);
                                                }
                            }
                        }
                    }
                    // target rows: degree am + row, row = 16 t + row(lk, r)
    #pragma unroll
                    for (int t = 0; t < kMaxTiles; ++t)
                        if (t < nt) {
                            const int nr0 = am + t * 16 + Mma<T>::row(lk, 0);
                            const int q0 = __umul24(nr0, nr0 + 1);
    #pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                // (nr0 + s)(nr0 + s + 1) = q0 + 2 s nr0 + s^2 + s,  s = r RSTEP
                                const int s = r * RSTEP;
                                const int q = q0 + __umul24(2 * s, nr0) + s * s + s;
                                const bool ok = t * 16 + Mma<T>::row(lk, r) < lent;
    #pragma unroll
                                for (int sg = 0; sg < 2; ++sg)
                                    if (sg < ns) {
                                        const int dst = __umul24(ok ? (sg ? q - am : q + am) : spare, LDX);
    #pragma unroll
                                        for (int ct = 0; ct < NCT; ++ct) X[dst + ct * 16 + li] = c4[sg][t][ct][r];
                                    }
                            }
                        }
                }
            }
        }

        // stage 3: back rotation of both blocks of degree n, output rows indexed by the
        // order m = 0..n in both (so u'_m and v'_m of one m sit in the same lane and
        // register), times e^{-i m phi} of each column tile's direction:
        //   U += cos(m phi) u' - i sin(m phi) v',   V += cos(m phi) v' - i sin(m phi) u'
        // (the phase applied to x_m and x_{-m}, written in the (u, v) basis), summed over
        // the group's directions into the accumulators; unit = (n, row tile)
        if constexpr (PF) {
            int n3 = 0;   // the wave's next unit
            while (n3 < MAXU && skip(__builtin_amdgcn_readlane(vt3, n3))) ++n3;
            V4 anu = V4{0, 0, 0, 0}, anv = V4{0, 0, 0, 0};
            auto pre = [&](int uu) {   // the u block's fragment, then the v block's (one tile each)
                const int t = __builtin_amdgcn_readlane(vt3, uu);
                const T* Au = R3 + (t >> 9) + lane_o * 4;
                anu = Mma<T>::load4(Au);
                if ((t & 255) > 0) anv = Mma<T>::load4(Au + 256);
            };
            if (n3 < MAXU) pre(n3);
            __syncthreads();
#pragma unroll
            for (int uu = 0; uu < MAXU; ++uu) {
                if (uu == n3) {
                    // a packed unit (bit 8) holds degree n1 in rows 0..7 and n2 in rows 8..15;
                    // its u inputs are n1's n1+1 then n2's n2+1, its v inputs n1's n1 then
                    // n2's n2, and its rows take the phases of the orders row mod 8
                    const int code3 = __builtin_amdgcn_readlane(vt3, uu) & 511;
                    const bool pk = code3 & 256;
                    const int n = code3 & 255, n1 = (code3 >> 4) & 15, n2 = code3 & 15;
                    const int nku = pk ? (n1 + n2 + 5) >> 2 : (n + 4) >> 2;
                    const int nkv = pk ? (n1 + n2 + 3) >> 2 : (n + 3) >> 2;
                    const C2* psu = ps_lane - (pk ? 8 * (lk >> 1) : 0);
                    const V4 avu = anu, avv = anv;
                    int nx = uu + 1;
                    while (nx < MAXU && skip(__builtin_amdgcn_readlane(vt3, nx))) ++nx;
                    n3 = nx;
                    if (nx < MAXU) pre(nx);
                    V4 cu[NCT], cv[NCT];
#pragma unroll
                    for (int ct = 0; ct < NCT; ++ct) {
                        cu[ct] = V4{0, 0, 0, 0};
                        cv[ct] = V4{0, 0, 0, 0};
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int k = 4 * j + lk;   // past the inputs: A is 0, the rows finite
                        const int ru = pk ? (k <= n1 ? n1 * n1 + n1 + k : n2 * n2 + n2 + k - n1 - 1) : n * n + n + k;
                        const int rv = pk ? (k < n1 ? n1 * n1 + k : n2 * n2 + k - n1) : n * n + k;
                        T bu[NCT], bw[NCT];
#pragma unroll
                        for (int ct = 0; ct < NCT; ++ct) {
                            bu[ct] = X[ru * LDX + ct * 16 + li];
                            bw[ct] = X[rv * LDX + ct * 16 + li];
                        }
                        if (j < nku)
#pragma unroll
                            for (int ct = 0; ct < NCT; ++ct) cu[ct] = Mma<T>::mma(avu[j], bu[ct], cu[ct]);
                        if (j < nkv)
#pragma unroll
                            for (int ct = 0; ct < NCT; ++ct) cv[ct] = Mma<T>::mma(avv[j], bw[ct], cv[ct]);
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
#pragma unroll
                        for (int g = 0; g < G; ++g) {
                            const C2 pm = psu[g * TL + r * RSTEP];
                            if constexpr (SPL) {
                                // tiles (re, im) of the direction: U += c u' - i s v', V likewise
                                const T ur = cu[2 * g][r], ui = cu[2 * g + 1][r];
                                const T vr = cv[2 * g][r], vi = cv[2 * g + 1][r];
                                accu[uu][0][r] = fma(ur, pm.x, fma(vi, pm.y, accu[uu][0][r]));
                                accu[uu][1][r] = fma(ui, pm.x, fma(-vr, pm.y, accu[uu][1][r]));
                                accv[uu][0][r] = fma(vr, pm.x, fma(ui, pm.y, accv[uu][0][r]));
                                accv[uu][1][r] = fma(vi, pm.x, fma(-ur, pm.y, accv[uu][1][r]));
                            } else {
#pragma unroll
                            for (int cd = 0; cd < NCTD; ++cd) {
                                const T wu = cu[g * NCTD + cd][r], wv = cv[g * NCTD + cd][r];
                                accu[uu][cd][r] = fma(wu, pm.x, fma(Mma<T>::swap_pair(wv), pm.y, accu[uu][cd][r]));
                                accv[uu][cd][r] = fma(wv, pm.x, fma(Mma<T>::swap_pair(wu), pm.y, accv[uu][cd][r]));
                            }
                            }
                        }
                    }
                }
            }
        } else {
            __syncthreads();
#pragma unroll
            for (int uu = 0; uu < MAXU; ++uu) {
                const int u = wave + 4 * uu;
                if (u < a.n_units3 && (a.units3[u] & 255) <= pt) {
                    const int code = a.units3[u];
                    const int n = code & 255, mt = code >> 8;
                    const int ntu = m2l_tiles(n + 1), nku = (n + 4) >> 2;   // u: n+1 inputs
                    const int ntv = m2l_tiles(n), nkv = (n + 3) >> 2;       // v: n inputs
                    const T* Au = R3 + a.rot_off3[2 * n] + (size_t)mt * ntu * 256;
                    const T* Av = R3 + a.rot_off3[2 * n + 1] + (size_t)mt * ntv * 256;
                    V4 cu[NCT], cv[NCT];
#pragma unroll
                    for (int ct = 0; ct < NCT; ++ct) {
                        cu[ct] = V4{0, 0, 0, 0};
                        cv[ct] = V4{0, 0, 0, 0};
                    }
                    for (int kg = 0; kg < ntu; ++kg) {
                        const V4 av = Mma<T>::load4(Au + ((size_t)kg * 64 + lane_o) * 4);
                        T bv[4][NCT];
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int row = n * n + n + 16 * kg + 4 * j + lk;
#pragma unroll
                            for (int ct = 0; ct < NCT; ++ct) bv[j][ct] = X[row * LDX + ct * 16 + li];
                        }
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            if (4 * kg + j < nku)
#pragma unroll
                                for (int ct = 0; ct < NCT; ++ct) cu[ct] = Mma<T>::mma(av[j], bv[j][ct], cu[ct]);
                    }
                    for (int kg = 0; kg < ntv; ++kg) {
                        const V4 av = Mma<T>::load4(Av + ((size_t)kg * 64 + lane_o) * 4);
                        T bv[4][NCT];
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int row = n * n + 16 * kg + 4 * j + lk;
#pragma unroll
                            for (int ct = 0; ct < NCT; ++ct) bv[j][ct] = X[row * LDX + ct * 16 + li];
                        }
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            if (4 * kg + j < nkv)
#pragma unroll
                                for (int ct = 0; ct < NCT; ++ct) cv[ct] = Mma<T>::mma(av[j], bv[j][ct], cv[ct]);
                    }
                    // Re U += c Re u' + s Im v', Im U += c Im u' - s Re v' (V likewise with u, v
                    // exchanged): the partner lane holds the other half, the table the sine with
                    // this lane's sign.  Rows past the blocks are zero; their entries are finite.
                    const C2* pu = ps_lane + mt * 16;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
#pragma unroll
                        for (int g = 0; g < G; ++g) {
                            const C2 pm = pu[g * TL + r * RSTEP];
                            if constexpr (SPL) {
                                const T ur = cu[2 * g][r], ui = cu[2 * g + 1][r];
                                const T vr = cv[2 * g][r], vi = cv[2 * g + 1][r];
                                accu[uu][0][r] = fma(ur, pm.x, fma(vi, pm.y, accu[uu][0][r]));
                                accu[uu][1][r] = fma(ui, pm.x, fma(-vr, pm.y, accu[uu][1][r]));
                                accv[uu][0][r] = fma(vr, pm.x, fma(ui, pm.y, accv[uu][0][r]));
                                accv[uu][1][r] = fma(vi, pm.x, fma(-ur, pm.y, accv[uu][1][r]));
                            } else {
#pragma unroll
                            for (int cd = 0; cd < NCTD; ++cd) {
                                const T wu = cu[g * NCTD + cd][r], wv = cv[g * NCTD + cd][r];
                                accu[uu][cd][r] = fma(wu, pm.x, fma(Mma<T>::swap_pair(wv), pm.y, accu[uu][cd][r]));
                                accv[uu][cd][r] = fma(wv, pm.x, fma(Mma<T>::swap_pair(wu), pm.y, accv[uu][cd][r]));
                            }
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
    }

    // back to the x basis and added to the FP64 local expansions:
    // x_0 = U_0, x_m = (U + V) / 2, x_{-m} = (-1)^m (U - V) / 2
#pragma unroll
    for (int uu = 0; uu < MAXU; ++uu) {
        const int u = wave + 4 * uu;
        const int code = u < a.n_units3 ? a.units3[u] : 255;
        // a packed unit (FP32 one-tile instances): degree n1 in rows 0..7, n2 in rows 8..15
        const bool pk = PF && (code & 256);
        if (pk ? ((code >> 4) & 15) <= pt : (code & 255) <= pt) {
#pragma unroll
            for (int cd = 0; cd < NCTD; ++cd) {
                const int col = cd * 16 + li;
                const int t = tg[SPL ? li : col >> 1];
                if (t >= 0) {
                    double* dst = (double*)(a.L + a.exp_off[t]) + (SPL ? cd : col & 1);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = Mma<T>::row(lk, r);
                        const int n = pk ? (row < 8 ? (code >> 4) & 15 : code & 15) : code & 255;
                        const int m = pk ? row & 7 : (code >> 8) * 16 + row;
                        if (m <= n && n <= pt) {
                            const double U = (double)accu[uu][cd][r], V = (double)accv[uu][cd][r];
                            const double xp = m == 0 ? U : 0.5 * (U + V);
                            const double xm = (m & 1) ? -0.5 * (U - V) : 0.5 * (U - V);
                            double* yp = dst + 2 * (n * n + n + m);
                            double* ym = dst + 2 * (n * n + n - m);
                            if (item.w) {
                                atomicAdd(yp, xp);
                                if (m > 0) atomicAdd(ym, xm);
                            } else if (a.overwrite) {
                                *yp = xp;
                                if (m > 0) *ym = xm;
                            } else {
                                *yp += xp;
                                if (m > 0) *ym += xm;
                            }
                        }
                    }
                }
            }
        }
    }
}

// One launch per translation: workgroup i runs on XCD i % 8, and XCD x takes the contiguous
// item range [x n/8, (x+1) n/8) (the host pads the item list to a multiple of 8 and orders
// it in space, so each XCD's L2 keeps one region's multipoles).
template <typename T, int TB, int G, int KT, int MAXU>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(m2l_block_min_waves<T, TB, G, KT>())))
void m2l_block_kernel(M2LBlockArgs a) {
    const int per_xcd = a.n_items / kXcds;
    m2l_block_body<T, TB, G, KT, MAXU>(a, a.items[(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds]);
}

// Persistent form for batches: resident workgroups take (translation, item) work from the
// queues q (one per XCD, spatially ordered, stealing from the others when their own runs
// dry) until all are empty; args[w.x] describes translation w.x.
// The loop keeps a few lane values live across items; FP32 instances with one tile per
// block are held to the 4 waves per SIMD of the one-launch form.
template <typename T, int TB, int G, int KT>
constexpr int m2l_persistent_min_waves() {
    return sizeof(T) == 4 && KT == 1 && TB * G <= 32 ? 4 : m2l_block_min_waves<T, TB, G, KT>();
}

template <typename T, int TB, int G, int KT, int MAXU>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(m2l_persistent_min_waves<T, TB, G, KT>())))
void m2l_block_persistent(const M2LBlockArgs* __restrict__ args, WorkQueues q) {
    __shared__ int s_w;
    const int home = blockIdx.x % kXcds;
    for (;;) {
        if (threadIdx.x == 0) s_w = work_dequeue(q, home);
        __syncthreads();
        // uniform across the workgroup: the descriptor is read with scalar loads
        const int w = __builtin_amdgcn_readfirstlane(s_w);
        __syncthreads();
        if (w < 0) break;
        const int2 wk = q.work[w];
        const M2LBlockArgs a = args[__builtin_amdgcn_readfirstlane(wk.x)];
        m2l_block_body<T, TB, G, KT, MAXU>(a, a.items[__builtin_amdgcn_readfirstlane(wk.y)]);
        __syncthreads();
    }
}

__global__ void to_float_kernel(const double2* in, float2* out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = make_float2((float)in[i].x, (float)in[i].y);
}

__global__ void to_c64_kernel(const double2* in, float2* out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float2((float)in[i].x, (float)in[i].y);
}

// a launch of the persistent form: device args, queues and the resident grid
struct PersistentLaunch {
    const M2LBlockArgs* d_args;
    WorkQueues q;
    int grid;            // 0: return the occupancy-limited resident grid in *grid_out
    int* grid_out;
};

template <typename T, int TB, int G, int KT, int MAXU>
void launch_one(const M2LBlockArgs* a, size_t shm, hipStream_t s, const PersistentLaunch* pl = nullptr) {
    if (pl) {
        const void* fn = (const void*)m2l_block_persistent<T, TB, G, KT, MAXU>;
        if (pl->grid == 0) {
            int dev = 0, cus = 0, per_cu = 0;
            (void)hipGetDevice(&dev);
            // the launch's own dynamic LDS (the kernel also holds a few static bytes)
            if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm) != hipSuccess)
                fail(-3, "hipFuncSetAttribute(m2l persistent) failed");
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, shm) != hipSuccess)
                fail(-3, "occupancy query (m2l persistent) failed");
            *pl->grid_out = std::max(kXcds, (std::max(1, per_cu) * cus) / kXcds * kXcds);
            return;
        }
        hipLaunchKernelGGL((m2l_block_persistent<T, TB, G, KT, MAXU>), dim3(pl->grid), dim3(256), shm, s,
                           pl->d_args, pl->q);
        return;
    }
    if (!a) {   // prepare: allow the dynamic LDS this instantiation may use
        int dev = 0, lds = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) lds = 65536;
        hipError_t e = hipFuncSetAttribute((const void*)m2l_block_kernel<T, TB, G, KT, MAXU>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) fail(-3, std::string("hipFuncSetAttribute(m2l) failed: ") + hipGetErrorString(e));
        return;
    }
    hipLaunchKernelGGL((m2l_block_kernel<T, TB, G, KT, MAXU>), dim3(a->n_items), dim3(256), shm, s, *a);
}

// (KT, MAXU) follow from the table degree: KT = m2l_tiles(p + 1), MAXU = m2l_block_maxu(p),
// or m2l_block_maxu_packed(p) for the FP32 one-tile instances (the host packs their units)
template <typename T, int TB, int G>
void launch_p(const M2LBlockArgs* a, size_t shm, int p, hipStream_t s, const PersistentLaunch* pl) {
    constexpr bool PK = sizeof(T) == 4;
    if (p <= 7)
        launch_one<T, TB, G, 1, PK ? m2l_block_maxu_packed(7) : m2l_block_maxu(7)>(a, shm, s, pl);
    else if (p <= 11)
        launch_one<T, TB, G, 1, PK ? m2l_block_maxu_packed(11) : m2l_block_maxu(11)>(a, shm, s, pl);
    else if (p <= 15)
        launch_one<T, TB, G, 1, PK ? m2l_block_maxu_packed(15) : m2l_block_maxu(15)>(a, shm, s, pl);
    else if (p <= 17)
        launch_one<T, TB, G, 2, m2l_block_maxu(17)>(a, shm, s, pl);
    else if (p <= 19)
        launch_one<T, TB, G, 2, m2l_block_maxu(19)>(a, shm, s, pl);
    else if (p <= kM2LBlockMaxP)
        launch_one<T, TB, G, 2, m2l_block_maxu(23)>(a, shm, s, pl);
    else
        fail(-9, "internal: degree above the block kernel's range");
}

template <typename T, int TB>
void launch_g(const M2LBlockArgs* a, size_t shm, int g, int p, hipStream_t s, const PersistentLaunch* pl) {
    if (g == 1)
        launch_p<T, TB, 1>(a, shm, p, s, pl);
    else if (g == 2)
        launch_p<T, TB, 2>(a, shm, p, s, pl);
    else if (g == 4)
        launch_p<T, TB, 4>(a, shm, p, s, pl);
    else
        fail(-9, "internal: unsupported M2L group size");
}

void dispatch(const M2LBlockArgs* a, int p, int tb, int g, int fp32, hipStream_t s,
              const PersistentLaunch* pl = nullptr, size_t shm_min = 0) {
    const size_t shm = std::max(shm_min, m2l_block_lds_bytes(p, tb, g, fp32 ? 4 : 8));
    if (fp32) {
        if (tb == 16)
            launch_g<float, 16>(a, shm, g, p, s, pl);
        else if (tb == 8)
            launch_g<float, 8>(a, shm, g, p, s, pl);
        else
            fail(-9, "internal: unsupported M2L block size");
    } else {
        if (tb == 8)
            launch_g<double, 8>(a, shm, g, p, s, pl);
        else if (tb == 16 && g == 1)
            launch_p<double, 16, 1>(a, shm, p, s, pl);
        else
            fail(-9, "internal: unsupported M2L block size");
    }
}

}  // namespace

void prepare_m2l_block(int p, int tb, int g, int fp32) { dispatch(nullptr, p, tb, g, fp32, nullptr); }

void launch_m2l_block(const M2LBlockArgs& a, int tb, int g, int fp32, hipStream_t s) {
    if (a.n_items <= 0) return;
    dispatch(&a, std::max(a.p_tab, a.pt_tab), tb, g, fp32, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(-3, std::string("kernel launch failed (m2l_block): ") + hipGetErrorString(e));
}

int m2l_block_persistent_grid(int p, int tb, int g, int fp32, size_t shm) {
    int grid = 0;
    const PersistentLaunch pl{nullptr, WorkQueues{}, 0, &grid};
    dispatch(nullptr, p, tb, g, fp32, nullptr, &pl, shm);
    return grid;
}

void launch_m2l_block_persistent(const M2LBlockArgs* d_args, const WorkQueues& q, int grid, int p, int tb, int g,
                                 int fp32, size_t shm, hipStream_t s) {
    if (grid <= 0) return;
    const PersistentLaunch pl{d_args, q, grid, nullptr};
    dispatch(nullptr, p, tb, g, fp32, s, &pl, shm);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(-3, std::string("kernel launch failed (m2l_block persistent): ") + hipGetErrorString(e));
}

void launch_to_float(const double2* in, float2* out, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(to_float_kernel, dim3((unsigned)blocks), dim3(256), 0, s, in, out, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(-3, std::string("kernel launch failed (to_float): ") + hipGetErrorString(e));
}

__global__ void zero_boxes_kernel(const int* boxes, int n, const int64_t* exp_off, int ncoef, double2* E) {
    const int b = blockIdx.x;
    if (b >= n) return;
    double2* e = E + exp_off[boxes[b]];
    for (int c = threadIdx.x; c < ncoef; c += blockDim.x) e[c] = make_double2(0.0, 0.0);
}

void launch_zero_boxes(const int* boxes, int n, const int64_t* exp_off, int ncoef, double2* E, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(zero_boxes_kernel, dim3(n), dim3(256), 0, s, boxes, n, exp_off, ncoef, E);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(-3, std::string("kernel launch failed (zero_boxes): ") + hipGetErrorString(e));
}

void launch_to_c64(const double2* in, float2* out, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(to_c64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(-3, std::string("kernel launch failed (to_c64): ") + hipGetErrorString(e));
}

}  // namespace hfmm
