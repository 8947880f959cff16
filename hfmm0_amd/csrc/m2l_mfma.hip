// M2L on matrix cores: all pairs of a block of target boxes that share a translation
// operator are contracted together.
//
// A workgroup (4 waves) owns TB target boxes of one level that have the same position
// (octant) inside their parents, so their V lists use the same set of displacements.
// For each displacement d present in the block it
//   0. gathers the TB source multipoles into LDS, multiplied by e^{i m phi_d}
//      (columns = (pair, re/im), rows = coefficient n*n+n+m);
//   1. rotates each degree block in place: X_n <- E_n(theta_d)^T X_n   (real x complex)
//   2. applies the coaxial operator per order m in place: X^m <- T^{|m|}(|d|) X^m, as
//      Re(T) * X + Im(T) * (X with re/im columns swapped and signed);
//   3. rotates back, W_n = E_n X_n, multiplies by e^{-i m phi_d} (the partner lane holds
//      the other half of the complex value) and accumulates in registers.
// Every product is a 16x16x4 MFMA tile.  The A operands come from operator tables that
// the host stores in fragment order, zero-padded to whole 16x16 tiles, so a lane loads
// the A values of four k-steps with one 16-byte (FP32) or two 16-byte (FP64) loads and
// no bounds test.  B operands are read from LDS with row indices clamped into the block
// (their products with the zero padding of A vanish), and each A fragment is applied to
// every 16-column tile of the block.  The stage-3 accumulators stay in registers across
// all displacements and are added to the FP64 local expansions once at the end.
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"

namespace hfmm {

// LDS: the expansion block (with its scratch row), then p+1 complex phases
size_t m2l_block_lds_bytes(int p, int tb, int precision_bytes) {
    return ((size_t)m2l_lds_x_elems(p, tb) + 2 * (size_t)(p + 1)) * precision_bytes;
}

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
    // output element r of lane group lk (= lane / 16) is row 4 lk + r of the 16x16 tile
    __device__ static inline int row(int lk, int r) { return 4 * lk + r; }
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
    __device__ static inline int row(int lk, int r) { return lk + 4 * r; }
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

__device__ inline void decode_nm(int c, int& n, int& m) {
    n = (int)sqrtf((float)c + 0.5f);
    if ((n + 1) * (n + 1) <= c) ++n;
    if (n * n > c) --n;
    m = c - n * n - n;
}

constexpr int kMaxTiles = (2 * kM2LBlockMaxP + 1 + 15) / 16;   // row tiles of a degree block
constexpr int kMaxTiles2 = (kM2LBlockMaxP + 1 + 15) / 16;      // row tiles of an order block

// Work item: entries [e0, e1) of one target block.  When a block is split over several
// items (coarse levels, to fill the GPU) the results are added atomically.
template <typename T, int TB, int MAXU>
__global__ __launch_bounds__(256) void m2l_block_kernel(M2LBlockArgs a) {
    using V4 = typename Mma<T>::v4;
    using C2 = typename Mma<T>::c2;
    constexpr int NCOL = 2 * TB;            // (pair, re/im) columns
    constexpr int LDX = NCOL + 1;           // odd row stride in LDS
    constexpr int NCT = NCOL / 16;          // 16-wide column tiles
    constexpr int NG = 8;                   // gathered values per thread and batch
    extern __shared__ char smem_raw[];
    const int p = a.p, P1 = p + 1, nc = P1 * P1;
    const int spare = (a.p_tab + 1) * (a.p_tab + 1);   // scratch row for out-of-block stores
    T* X = (T*)smem_raw;
    C2* Ps = (C2*)(X + m2l_lds_x_elems(a.p_tab, TB));
    // the wave index is uniform across a wavefront: keep it (and all unit bookkeeping
    // derived from it) in scalar registers so that the unit branches are scalar
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int li = lane & 15, lk = lane >> 4;
    const int4 item = a.items[blockIdx.x];  // (block, e0, e1, atomic)
    const int* tg = a.blk_targets + (size_t)item.x * TB;
    const C2* Mexp = (const C2*)a.M;

    V4 acc[MAXU][NCT];
#pragma unroll
    for (int u = 0; u < MAXU; ++u)
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) acc[u][ct] = V4{0, 0, 0, 0};

    for (int e = item.y; e < item.z; ++e) {
        const int d = a.ent_dir[e];
        const DispInfo di = a.dinfo[d];
        const T* R1 = (const T*)a.rot1 + (size_t)di.theta * a.rot_stride;
        const T* R3 = (const T*)a.rot3 + (size_t)di.theta * a.rot_stride;
        const T* Tz = (const T*)a.coax + (size_t)di.dist * a.coax_stride;
        const C2* phg = (const C2*)a.phase + (size_t)d * a.phase_stride;
        const int* soff = a.ent_off + (size_t)e * TB;

        // stage 0: gather + e^{i m phi}; the loads of a batch are issued before any is used
        if (threadIdx.x < P1) Ps[threadIdx.x] = phg[threadIdx.x];
        for (int base = 0; base < TB * nc; base += 256 * NG) {
            C2 raw[NG];
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const int idx = base + threadIdx.x + 256 * g;
                raw[g].x = 0;
                raw[g].y = 0;
                if (idx < TB * nc) {
                    const int j = idx / nc, c = idx - j * nc;
                    const int o = soff[j];
                    if (o >= 0) raw[g] = Mexp[(int64_t)o + c];
                }
            }
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const int idx = base + threadIdx.x + 256 * g;
                if (idx < TB * nc) {
                    const int j = idx / nc, c = idx - j * nc;
                    int n, m;
                    decode_nm(c, n, m);
                    C2 pm = phg[m < 0 ? -m : m];
                    if (m < 0) pm.y = -pm.y;
                    X[c * LDX + 2 * j] = raw[g].x * pm.x - raw[g].y * pm.y;
                    X[c * LDX + 2 * j + 1] = raw[g].x * pm.y + raw[g].y * pm.x;
                }
            }
        }
        __syncthreads();

        // stage 1: X_n <- E_n^T X_n in place.  A unit is one degree n with all its row and
        // column tiles: the wave writes back only after all products are formed, and no
        // other wave touches the rows of degree n.
        for (int u = wave; u < a.n_units1; u += 4) {
            const int n = a.units1[u];
            if (n > p) continue;
            const int w = 2 * n + 1, base = n * n;
            const int nt = m2l_tiles(w), nks = (w + 3) >> 2;
            const T* A = R1 + a.rot_off[n];
            V4 c4[kMaxTiles][NCT];
#pragma unroll
            for (int mt = 0; mt < kMaxTiles; ++mt)
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct) c4[mt][ct] = V4{0, 0, 0, 0};
            for (int kg = 0; kg < nt; ++kg) {
                V4 av[kMaxTiles];
#pragma unroll
                for (int mt = 0; mt < kMaxTiles; ++mt)
                    if (mt < nt) av[mt] = Mma<T>::load4(A + ((size_t)(mt * nt + kg) * 64 + lane) * 4);
                T bv[4][NCT];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int row = base + min(16 * kg + 4 * j + lk, w - 1);
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
                        const int dst = (row < w ? base + row : spare) * LDX;
#pragma unroll
                        for (int ct = 0; ct < NCT; ++ct) X[dst + ct * 16 + li] = c4[mt][ct][r];
                    }
        }
        __syncthreads();

        // stage 2: X^m <- T^{|m|} X^m in place, one order m (all tiles) per unit.  The
        // table blocks are built for p_tab; source degrees beyond the evaluated p are
        // read as zero.
        for (int u = wave; u < a.n_units2; u += 4) {
            const int m = a.units2[u] - a.p_tab;
            const int am = m < 0 ? -m : m;
            if (am > p) continue;
            const int len = P1 - am;
            const int ntab = m2l_tiles(a.p_tab + 1 - am);   // tiles of the table block
            const int nt = m2l_tiles(len), nks = (len + 3) >> 2;
            const T* A = Tz + a.coax_off[am];
            V4 c4[kMaxTiles2][NCT];
#pragma unroll
            for (int t = 0; t < kMaxTiles2; ++t)
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct) c4[t][ct] = V4{0, 0, 0, 0};
            for (int kg = 0; kg < nt; ++kg) {
                V4 are[kMaxTiles2], aim[kMaxTiles2];
#pragma unroll
                for (int t = 0; t < kMaxTiles2; ++t)
                    if (t < nt) {
                        const T* At = A + ((size_t)(t * ntab + kg) * 64 + lane) * 8;
                        are[t] = Mma<T>::load4(At);
                        aim[t] = Mma<T>::load4(At + 4);
                    }
                T b0[4][NCT], b1[4][NCT];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int kin = 16 * kg + 4 * j + lk;
                    const bool kv = kin < len;
                    const int nn = am + min(kin, len - 1);
                    const int cin = nn * nn + nn + m;
#pragma unroll
                    for (int ct = 0; ct < NCT; ++ct) {
                        const int col = ct * 16 + li;
                        const T v0 = X[cin * LDX + col], v1 = X[cin * LDX + (col ^ 1)];
                        b0[j][ct] = kv ? v0 : (T)0;
                        b1[j][ct] = kv ? ((col & 1) ? v1 : -v1) : (T)0;
                    }
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (4 * kg + j < nks)
#pragma unroll
                        for (int t = 0; t < kMaxTiles2; ++t)
                            if (t < nt)
#pragma unroll
                                for (int ct = 0; ct < NCT; ++ct) {
                                    c4[t][ct] = Mma<T>::mma(are[t][j], b0[j][ct], c4[t][ct]);
                                    c4[t][ct] = Mma<T>::mma(aim[t][j], b1[j][ct], c4[t][ct]);
                                }
            }
#pragma unroll
            for (int t = 0; t < kMaxTiles2; ++t)
                if (t < nt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = t * 16 + Mma<T>::row(lk, r);
                        const int nn = am + row;
                        const int dst = (row < len ? nn * nn + nn + m : spare) * LDX;
#pragma unroll
                        for (int ct = 0; ct < NCT; ++ct) X[dst + ct * 16 + li] = c4[t][ct][r];
                    }
        }
        __syncthreads();

        // stage 3: W_n = E_n X_n, times e^{-i m phi}, accumulated; unit = (n, row tile)
#pragma unroll
        for (int uu = 0; uu < MAXU; ++uu) {
            const int u = wave + 4 * uu;
            if (u < a.n_units3 && (a.units3[u] & 255) <= p) {
                const int code = a.units3[u];
                const int n = code & 255, mt = code >> 8;
                const int w = 2 * n + 1, base = n * n;
                const int nt = m2l_tiles(w), nks = (w + 3) >> 2;
                const T* A = R3 + a.rot_off[n] + (size_t)mt * nt * 256;
                V4 c4[NCT];
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct) c4[ct] = V4{0, 0, 0, 0};
                for (int kg = 0; kg < nt; ++kg) {
                    const V4 av = Mma<T>::load4(A + ((size_t)kg * 64 + lane) * 4);
                    T bv[4][NCT];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int row = base + min(16 * kg + 4 * j + lk, w - 1);
#pragma unroll
                        for (int ct = 0; ct < NCT; ++ct) bv[j][ct] = X[row * LDX + ct * 16 + li];
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (4 * kg + j < nks)
#pragma unroll
                            for (int ct = 0; ct < NCT; ++ct) c4[ct] = Mma<T>::mma(av[j], bv[j][ct], c4[ct]);
                }
                // rows past the block are zero (A is zero-padded); their phase is irrelevant
                const bool im_part = (li & 1) != 0;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = mt * 16 + Mma<T>::row(lk, r) - n;
                    const int am = m < 0 ? -m : m;
                    const C2 pm = Ps[min(am, p)];
                    const T cs = pm.x, sn = m < 0 ? -pm.y : pm.y;
#pragma unroll
                    for (int ct = 0; ct < NCT; ++ct) {
                        const T v = c4[ct][r];
                        const T partner = Mma<T>::swap_pair(v);
                        acc[uu][ct][r] += im_part ? (v * cs - partner * sn) : (v * cs + partner * sn);
                    }
                }
            }
        }
        __syncthreads();
    }

    // add to the FP64 local expansions
#pragma unroll
    for (int uu = 0; uu < MAXU; ++uu) {
        const int u = wave + 4 * uu;
        if (u < a.n_units3 && (a.units3[u] & 255) <= p) {
            const int code = a.units3[u];
            const int n = code & 255, mt = code >> 8;
#pragma unroll
            for (int ct = 0; ct < NCT; ++ct) {
                const int col = ct * 16 + li;
                const int t = tg[col >> 1];
                if (t >= 0) {
                    double* dst = (double*)(a.L + a.exp_off[t]) + (col & 1);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = mt * 16 + Mma<T>::row(lk, r);
                        if (row < 2 * n + 1) {
                            double* x = dst + 2 * (n * n + row);
                            if (item.w)
                                atomicAdd(x, (double)acc[uu][ct][r]);
                            else
                                *x += (double)acc[uu][ct][r];
                        }
                    }
                }
            }
        }
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

template <typename T, int TB, int MAXU>
void launch_one(const M2LBlockArgs* a, size_t shm, hipStream_t s) {
    if (!a) {   // prepare: allow the dynamic LDS this instantiation may use
        int dev = 0, lds = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) lds = 65536;
        hipError_t e = hipFuncSetAttribute((const void*)m2l_block_kernel<T, TB, MAXU>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) fail(-3, std::string("hipFuncSetAttribute(m2l) failed: ") + hipGetErrorString(e));
        return;
    }
    hipLaunchKernelGGL((m2l_block_kernel<T, TB, MAXU>), dim3(a->n_items), dim3(256), shm, s, *a);
}

template <typename T, int TB, int MAXU>
void launch_checked(const M2LBlockArgs* a, size_t shm, hipStream_t s) {
    if constexpr (m2l_acc_regs(MAXU, TB, sizeof(T)) <= kM2LMaxAccRegs)
        launch_one<T, TB, MAXU>(a, shm, s);
    else
        fail(-9, "internal: M2L block configuration exceeds the register budget");
}

template <typename T, int TB>
void launch_tb(const M2LBlockArgs* a, size_t shm, int maxu, hipStream_t s) {
    if (maxu <= 4)
        launch_checked<T, TB, 4>(a, shm, s);
    else if (maxu <= 8)
        launch_checked<T, TB, 8>(a, shm, s);
    else if (maxu <= 12)
        launch_checked<T, TB, 12>(a, shm, s);
    else if (maxu <= 16)
        launch_checked<T, TB, 16>(a, shm, s);
    else
        fail(-9, "internal: too many M2L units per wave");
}

void dispatch(const M2LBlockArgs* a, int p, int n_units3, int tb, int fp32, hipStream_t s) {
    const int maxu = (n_units3 + 3) / 4;
    const size_t shm = m2l_block_lds_bytes(p, tb, fp32 ? 4 : 8);
    if (fp32) {
        if (tb == 32)
            launch_tb<float, 32>(a, shm, maxu, s);
        else if (tb == 16)
            launch_tb<float, 16>(a, shm, maxu, s);
        else if (tb == 8)
            launch_tb<float, 8>(a, shm, maxu, s);
        else
            fail(-9, "internal: unsupported M2L block size");
    } else {
        if (tb == 16)
            launch_tb<double, 16>(a, shm, maxu, s);
        else if (tb == 8)
            launch_tb<double, 8>(a, shm, maxu, s);
        else
            fail(-9, "internal: unsupported M2L block size");
    }
}

}  // namespace

void prepare_m2l_block(int p, int n_units3, int tb, int fp32) { dispatch(nullptr, p, n_units3, tb, fp32, nullptr); }

void launch_m2l_block(const M2LBlockArgs& a, int tb, int fp32, hipStream_t s) {
    if (a.n_items <= 0) return;
    dispatch(&a, a.p_tab, a.n_units3, tb, fp32, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(-3, std::string("kernel launch failed (m2l_block): ") + hipGetErrorString(e));
}

void launch_to_float(const double2* in, float2* out, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(to_float_kernel, dim3((unsigned)blocks), dim3(256), 0, s, in, out, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(-3, std::string("kernel launch failed (to_float): ") + hipGetErrorString(e));
}

void launch_to_c64(const double2* in, float2* out, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(to_c64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(-3, std::string("kernel launch failed (to_c64): ") + hipGetErrorString(e));
}

}  // namespace hfmm
