// M2L on matrix cores: all pairs of a block of target boxes that share a translation
// operator are contracted together.
//
// A workgroup (4 waves) owns TB target boxes of one level that have the same position
// (octant) inside their parents, so their V lists use the same set of displacements.
// For each displacement d present in the block it
//   0. gathers the TB source multipoles into LDS, multiplied by e^{i m phi_d}
//      (columns = (pair, re/im), rows = coefficient n*n+n+m);
//   1. rotates each degree block: Y_n = E_n(theta_d)^T X_n        (real x complex)
//   2. applies the coaxial operator per order m: Z^m = T^{|m|}(|d|) Y^m, as
//      Re(T) * Y + Im(T) * (Y with re/im columns swapped and signed);
//   3. rotates back, W_n = E_n Z_n, multiplies by e^{-i m phi_d} (partner lane holds the
//      other half of the complex value) and accumulates in registers.
// Every product is a 16x16x4 MFMA tile (A from the operator tables, B from LDS; the
// FP32 and FP64 instructions place output rows differently, see Mma<T>::row);
// the accumulators of stage 3 stay in VGPRs across all displacements and are added
// to the FP64 local expansions once at the end.
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"

namespace hfmm {

size_t m2l_block_lds_bytes(int p, int tb, int precision_bytes) {
    const size_t nc = (size_t)(p + 1) * (p + 1);
    return 2 * nc * (2 * tb + 1) * precision_bytes;
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
};

__device__ inline void decode_nm(int c, int& n, int& m) {
    n = (int)sqrtf((float)c + 0.5f);
    if ((n + 1) * (n + 1) <= c) ++n;
    if (n * n > c) --n;
    m = c - n * n - n;
}

// sum_{i=1}^{x} i^2
__device__ inline int sq_sum(int x) { return x * (x + 1) * (2 * x + 1) / 6; }

template <typename T, int TB, int MAXU>
__global__ __launch_bounds__(256) void m2l_block_kernel(M2LBlockArgs a) {
    using V4 = typename Mma<T>::v4;
    using C2 = typename Mma<T>::c2;
    constexpr int NCOL = 2 * TB;            // (pair, re/im) columns
    constexpr int LDX = NCOL + 1;           // odd row stride in LDS
    constexpr int NCT = (NCOL + 15) / 16;   // 16-wide column tiles
    extern __shared__ char smem_raw[];
    const int p = a.p, P1 = p + 1, nc = P1 * P1;
    T* X = (T*)smem_raw;
    T* Y = X + (size_t)nc * LDX;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int blk = blockIdx.x;
    const int* tg = a.blk_targets + (size_t)blk * TB;
    const C2* Mexp = (const C2*)a.M;

    V4 acc[MAXU];
#pragma unroll
    for (int u = 0; u < MAXU; ++u) acc[u] = V4{0, 0, 0, 0};

    for (int e = a.blk_ptr[blk]; e < a.blk_ptr[blk + 1]; ++e) {
        const int d = a.ent_dir[e];
        const DispInfo di = a.dinfo[d];
        const T* E = (const T*)a.rot + (size_t)di.theta * a.rot_stride;
        const C2* ph = (const C2*)a.phase + (size_t)d * a.phase_stride;
        const C2* Tz = (const C2*)a.coax + (size_t)di.dist * a.coax_stride;
        const int* srcs = a.ent_src + (size_t)e * TB;

        // stage 0: gather + e^{i m phi}
        for (int idx = threadIdx.x; idx < TB * nc; idx += 256) {
            const int j = idx / nc, c = idx - j * nc;
            const int s = srcs[j];
            T vr = 0, vi = 0;
            if (s >= 0) {
                const C2 mv = Mexp[a.exp_off[s] + c];
                int n, m;
                decode_nm(c, n, m);
                C2 pm = ph[m < 0 ? -m : m];
                if (m < 0) pm.y = -pm.y;
                vr = mv.x * pm.x - mv.y * pm.y;
                vi = mv.x * pm.y + mv.y * pm.x;
            }
            X[c * LDX + 2 * j] = vr;
            X[c * LDX + 2 * j + 1] = vi;
        }
        __syncthreads();

        // stage 1: Y_n = E_n^T X_n
        for (int u = wave; u < a.n_units1; u += 4) {
            const int code = a.units1[u];
            const int n = code & 255, mt = (code >> 8) & 255, ct = code >> 16;
            const int w = 2 * n + 1, base = n * n;
            const T* En = E + n * (2 * n - 1) * (2 * n + 1) / 3;
            const int mo = mt * 16 + li;
            const int col = ct * 16 + li;
            const bool colok = col < NCOL;
            V4 c4 = V4{0, 0, 0, 0};
            for (int k0 = 0; k0 < w; k0 += 4) {
                const int kin = k0 + lk;
                const T av = (mo < w && kin < w) ? En[kin * w + mo] : (T)0;
                const T bv = (kin < w && colok) ? X[(base + kin) * LDX + col] : (T)0;
                c4 = Mma<T>::mma(av, bv, c4);
            }
            if (colok) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = mt * 16 + Mma<T>::row(lk, r);
                    if (row < w) Y[(base + row) * LDX + col] = c4[r];
                }
            }
        }
        __syncthreads();

        // stage 2: Z^m = T^{|m|} Y^m  (written over X)
        for (int u = wave; u < a.n_units2; u += 4) {
            const int code = a.units2[u];
            const int m = (code & 255) - p, mt = (code >> 8) & 255, ct = code >> 16;
            const int am = m < 0 ? -m : m;
            const int len = P1 - am;
            const C2* Tm = Tz + (sq_sum(P1) - sq_sum(P1 - am));
            const int nrow = mt * 16 + li;
            const int col = ct * 16 + li;
            const bool colok = col < NCOL;
            const T sgn = (col & 1) ? (T)1 : (T)-1;
            V4 c4 = V4{0, 0, 0, 0};
            for (int k0 = 0; k0 < len; k0 += 4) {
                const int kin = k0 + lk;
                const bool kv = kin < len;
                C2 t;
                t.x = 0;
                t.y = 0;
                if (nrow < len && kv) t = Tm[nrow * len + kin];
                const int nn = am + kin;
                const int cin = nn * nn + nn + m;
                const T b0 = (kv && colok) ? Y[cin * LDX + col] : (T)0;
                const T b1 = (kv && colok) ? sgn * Y[cin * LDX + (col ^ 1)] : (T)0;
                c4 = Mma<T>::mma(t.x, b0, c4);
                c4 = Mma<T>::mma(t.y, b1, c4);
            }
            if (colok) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = mt * 16 + Mma<T>::row(lk, r);
                    if (row < len) {
                        const int nn = am + row;
                        X[(nn * nn + nn + m) * LDX + col] = c4[r];
                    }
                }
            }
        }
        __syncthreads();

        // stage 3: W_n = E_n Z_n, times e^{-i m phi}, accumulate
#pragma unroll
        for (int uu = 0; uu < MAXU; ++uu) {
            const int u = wave + 4 * uu;
            if (u < a.n_units1) {
                const int code = a.units1[u];
                const int n = code & 255, mt = (code >> 8) & 255, ct = code >> 16;
                const int w = 2 * n + 1, base = n * n;
                const T* En = E + n * (2 * n - 1) * (2 * n + 1) / 3;
                const int mo = mt * 16 + li;
                const int col = ct * 16 + li;
                const bool colok = col < NCOL;
                V4 c4 = V4{0, 0, 0, 0};
                for (int k0 = 0; k0 < w; k0 += 4) {
                    const int kin = k0 + lk;
                    const T av = (mo < w && kin < w) ? En[mo * w + kin] : (T)0;
                    const T bv = (kin < w && colok) ? X[(base + kin) * LDX + col] : (T)0;
                    c4 = Mma<T>::mma(av, bv, c4);
                }
                const bool im_part = (col & 1) != 0;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = mt * 16 + Mma<T>::row(lk, r);
                    const int m = row - n;
                    const int am = m < 0 ? -m : m;
                    C2 pm;
                    pm.x = 1;
                    pm.y = 0;
                    if (row < w) pm = ph[am];
                    const T cs = pm.x, sn = m < 0 ? -pm.y : pm.y;
                    const T v = c4[r];
                    const T partner = __shfl_xor(v, 1);
                    acc[uu][r] += im_part ? (v * cs - partner * sn) : (v * cs + partner * sn);
                }
            }
        }
        __syncthreads();
    }

    // add to the FP64 local expansions
#pragma unroll
    for (int uu = 0; uu < MAXU; ++uu) {
        const int u = wave + 4 * uu;
        if (u < a.n_units1) {
            const int code = a.units1[u];
            const int n = code & 255, mt = (code >> 8) & 255, ct = code >> 16;
            const int col = ct * 16 + li;
            if (col < NCOL) {
                const int j = col >> 1, ri = col & 1;
                const int t = tg[j];
                if (t >= 0) {
                    double* dst = (double*)(a.L + a.exp_off[t]);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = mt * 16 + Mma<T>::row(lk, r);
                        if (row < 2 * n + 1) dst[2 * (n * n + row) + ri] += (double)acc[uu][r];
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
    hipLaunchKernelGGL((m2l_block_kernel<T, TB, MAXU>), dim3(a->n_blocks), dim3(256), shm, s, *a);
}

template <typename T, int TB>
void launch_tb(const M2LBlockArgs* a, size_t shm, int maxu, hipStream_t s) {
    if (maxu <= 8)
        launch_one<T, TB, 8>(a, shm, s);
    else if (maxu <= 12)
        launch_one<T, TB, 12>(a, shm, s);
    else if (maxu <= 16)
        launch_one<T, TB, 16>(a, shm, s);
    else if (maxu <= 24)
        launch_one<T, TB, 24>(a, shm, s);
    else if (maxu <= 32)
        launch_one<T, TB, 32>(a, shm, s);
    else if (maxu <= 48)
        launch_one<T, TB, 48>(a, shm, s);
    else
        fail(-9, "internal: too many M2L units per wave");
}

void dispatch(const M2LBlockArgs* a, int p, int n_units1, int tb, int fp32, hipStream_t s) {
    const int maxu = (n_units1 + 3) / 4;
    const size_t shm = m2l_block_lds_bytes(p, tb, fp32 ? 4 : 8);
    if (fp32) {
        if (tb == 16)
            launch_tb<float, 16>(a, shm, maxu, s);
        else if (tb == 8)
            launch_tb<float, 8>(a, shm, maxu, s);
        else
            launch_tb<float, 4>(a, shm, maxu, s);
    } else {
        if (tb == 16)
            launch_tb<double, 16>(a, shm, maxu, s);
        else if (tb == 8)
            launch_tb<double, 8>(a, shm, maxu, s);
        else
            launch_tb<double, 4>(a, shm, maxu, s);
    }
}

}  // namespace

void prepare_m2l_block(int p, int n_units1, int tb, int fp32) { dispatch(nullptr, p, n_units1, tb, fp32, nullptr); }

void launch_m2l_block(const M2LBlockArgs& a, int tb, int fp32, hipStream_t s) {
    if (a.n_blocks <= 0) return;
    dispatch(&a, a.p, a.n_units1, tb, fp32, s);
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

}  // namespace hfmm
