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
//   3. rotates back, W_n = E_n X_n, multiplies by e^{-i m phi_d} (partner lane holds the
//      other half of the complex value) and accumulates in registers.
// Every product is a 16x16x4 MFMA tile (A from the operator tables, B from LDS; the
// FP32 and FP64 instructions place output rows differently, see Mma<T>::row);
// the accumulators of stage 3 stay in VGPRs across all displacements and are added
// to the FP64 local expansions once at the end.
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"

namespace hfmm {

size_t m2l_block_lds_bytes(int p, int tb, int precision_bytes, int ops) {
    const size_t P1 = p + 1, nc = P1 * P1;
    size_t b = nc * (2 * tb + 1);
    if (!ops) return b * precision_bytes;
    const size_t rot_n = P1 * (2 * P1 - 1) * (2 * P1 + 1) / 3;
    const size_t tsz = P1 * (P1 + 1) * (2 * P1 + 1) / 6;
    b = ((b + 3) & ~(size_t)3) + ((rot_n + 3) & ~(size_t)3);
    return b * precision_bytes + (tsz + P1) * 2 * precision_bytes;
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

constexpr int kBlockMaxP = 23;                          // largest degree of the block kernel
constexpr int kMaxTiles = (2 * kBlockMaxP + 1 + 15) / 16; // row tiles of a degree block
constexpr int kMaxTiles2 = (kBlockMaxP + 1 + 15) / 16;    // row tiles of an order block

// sum_{i=1}^{x} i^2
__device__ inline int sq_sum(int x) { return x * (x + 1) * (2 * x + 1) / 6; }

// Work item of the kernel: entries [e0, e1) of one target block.  When a block is split
// over several items (coarse levels, to fill the GPU) the results are added atomically.
// OPS: the rotation block set, the coaxial table and the phases of the current
// displacement are copied into LDS (reloaded only when theta or |d| change; entries are
// sorted by them) so that the MFMA loops read every operand from LDS.
template <typename T, int TB, int MAXU, bool OPS>
__global__ __launch_bounds__(256) void m2l_block_kernel(M2LBlockArgs a) {
    using V4 = typename Mma<T>::v4;
    using C2 = typename Mma<T>::c2;
    constexpr int NCOL = 2 * TB;            // (pair, re/im) columns
    constexpr int LDX = NCOL + 1;           // odd row stride in LDS
    constexpr int KC = 8;                   // k-steps whose operands are loaded together
    constexpr int NG = 16;                  // gathered values per thread and batch
    extern __shared__ char smem_raw[];
    const int p = a.p, P1 = p + 1, nc = P1 * P1;
    const int PT1 = a.p_tab + 1;            // coaxial tables are packed for degree p_tab
    T* X = (T*)smem_raw;
    const int rot_n = P1 * (2 * P1 - 1) * (2 * P1 + 1) / 3;      // rotation blocks n <= p
    const int tsz = sq_sum(PT1);                                   // coaxial table entries
    T* Es = X + (((size_t)nc * LDX + 3) & ~(size_t)3);            // OPS only, 16-byte aligned
    C2* Ts = (C2*)(Es + ((rot_n + 3) & ~3));
    C2* Ps = Ts + tsz;
    int cur_theta = -1, cur_dist = -1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int4 item = a.items[blockIdx.x];  // (block, e0, e1, atomic)
    const int blk = item.x;
    const int* tg = a.blk_targets + (size_t)blk * TB;
    const C2* Mexp = (const C2*)a.M;

    V4 acc[MAXU];
#pragma unroll
    for (int u = 0; u < MAXU; ++u) acc[u] = V4{0, 0, 0, 0};

    for (int e = item.y; e < item.z; ++e) {
        const int d = a.ent_dir[e];
        const DispInfo di = a.dinfo[d];
        const T* Etg = (const T*)a.rot_t + (size_t)di.theta * a.rot_stride;  // transposed blocks
        const T* Eg = (const T*)a.rot + (size_t)di.theta * a.rot_stride;
        const C2* phg = (const C2*)a.phase + (size_t)d * a.phase_stride;
        const C2* Tg = (const C2*)a.coax + (size_t)di.dist * a.coax_stride;
        const int* soff = a.ent_off + (size_t)e * TB;
        const T* E = OPS ? Es : Eg;
        const C2* Tz = OPS ? Ts : Tg;
        const C2* ph = OPS ? Ps : phg;
        if (OPS) {
            if (di.theta != cur_theta) {
                for (int base = 0; base < rot_n; base += 256 * 8) {
                    T v[8];
#pragma unroll
                    for (int g = 0; g < 8; ++g) {
                        const int i = base + threadIdx.x + 256 * g;
                        v[g] = i < rot_n ? Eg[i] : (T)0;
                    }
#pragma unroll
                    for (int g = 0; g < 8; ++g) {
                        const int i = base + threadIdx.x + 256 * g;
                        if (i < rot_n) Es[i] = v[g];
                    }
                }
                cur_theta = di.theta;
            }
            if (di.dist != cur_dist) {
                for (int base = 0; base < tsz; base += 256 * 4) {
                    C2 v[4];
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        const int i = base + threadIdx.x + 256 * g;
                        if (i < tsz) v[g] = Tg[i];
                    }
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        const int i = base + threadIdx.x + 256 * g;
                        if (i < tsz) Ts[i] = v[g];
                    }
                }
                cur_dist = di.dist;
            }
            if (threadIdx.x < P1) Ps[threadIdx.x] = phg[threadIdx.x];
        }

        // stage 0: gather + e^{i m phi}; loads of a batch are issued before any is used
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

        // stage 1: X_n <- E_n^T X_n in place.  A unit is one degree n and one 16-column
        // group; the wave computes all its row tiles before writing them back, and no other
        // wave touches those rows and columns.  A[i][k] = E[k][i]: rows of E are contiguous.
        for (int u = wave; u < a.n_units1; u += 4) {
            const int code = a.units1[u];
            const int n = code & 255, ct = code >> 16;
            if (n > p) continue;
            const int w = 2 * n + 1, base = n * n;
            const int ntile = (w + 15) >> 4;
            const T* En = E + n * (2 * n - 1) * (2 * n + 1) / 3;
            const int col = ct * 16 + li;
            const bool colok = col < NCOL;
            V4 c4[kMaxTiles];
#pragma unroll
            for (int mt = 0; mt < kMaxTiles; ++mt) c4[mt] = V4{0, 0, 0, 0};
            for (int k0 = 0; k0 < w; k0 += 4 * KC) {
                T bv[KC];
#pragma unroll
                for (int q = 0; q < KC; ++q) {
                    const int kin = k0 + 4 * q + lk;
                    bv[q] = (kin < w && colok) ? X[(base + kin) * LDX + col] : (T)0;
                }
#pragma unroll
                for (int mt = 0; mt < kMaxTiles; ++mt) {
                    if (mt < ntile) {
                        const int mo = mt * 16 + li;
                        T av[KC];
#pragma unroll
                        for (int q = 0; q < KC; ++q) {
                            const int kin = k0 + 4 * q + lk;
                            av[q] = (mo < w && kin < w) ? En[kin * w + mo] : (T)0;
                        }
#pragma unroll
                        for (int q = 0; q < KC; ++q)
                            if (k0 + 4 * q < w) c4[mt] = Mma<T>::mma(av[q], bv[q], c4[mt]);
                    }
                }
            }
            if (colok) {
#pragma unroll
                for (int mt = 0; mt < kMaxTiles; ++mt)
                    if (mt < ntile) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int row = mt * 16 + Mma<T>::row(lk, r);
                            if (row < w) X[(base + row) * LDX + col] = c4[mt][r];
                        }
                    }
            }
        }
        __syncthreads();

        // stage 2: X^m <- T^{|m|} X^m in place, one order m and one column group per unit;
        // A[i][k] = T[n][n'] from the transposed table (16 rows of a k-step contiguous)
        for (int u = wave; u < a.n_units2; u += 4) {
            const int code = a.units2[u];
            const int m = (code & 255) - a.p_tab, ct = code >> 16;
            const int am = m < 0 ? -m : m;
            if (am > p) continue;
            const int len = P1 - am;
            const int lent = PT1 - am;          // row length of the packed (transposed) block
            const int ntile = (len + 15) >> 4;
            const C2* Tm = Tz + (sq_sum(PT1) - sq_sum(PT1 - am));
            const int col = ct * 16 + li;
            const bool colok = col < NCOL;
            const T sgn = (col & 1) ? (T)1 : (T)-1;
            V4 c4[kMaxTiles2];
#pragma unroll
            for (int nt = 0; nt < kMaxTiles2; ++nt) c4[nt] = V4{0, 0, 0, 0};
            for (int k0 = 0; k0 < len; k0 += 4 * KC) {
                T b0[KC], b1[KC];
#pragma unroll
                for (int q = 0; q < KC; ++q) {
                    const int kin = k0 + 4 * q + lk;
                    const bool kv = kin < len;
                    const int nn = am + kin;
                    const int cin = nn * nn + nn + m;
                    b0[q] = (kv && colok) ? X[cin * LDX + col] : (T)0;
                    b1[q] = (kv && colok) ? sgn * X[cin * LDX + (col ^ 1)] : (T)0;
                }
#pragma unroll
                for (int nt = 0; nt < kMaxTiles2; ++nt) {
                    if (nt < ntile) {
                        const int nrow = nt * 16 + li;
                        C2 tv[KC];
#pragma unroll
                        for (int q = 0; q < KC; ++q) {
                            const int kin = k0 + 4 * q + lk;
                            tv[q].x = 0;
                            tv[q].y = 0;
                            if (nrow < len && kin < len) tv[q] = Tm[kin * lent + nrow];
                        }
#pragma unroll
                        for (int q = 0; q < KC; ++q)
                            if (k0 + 4 * q < len) {
                                c4[nt] = Mma<T>::mma(tv[q].x, b0[q], c4[nt]);
                                c4[nt] = Mma<T>::mma(tv[q].y, b1[q], c4[nt]);
                            }
                    }
                }
            }
            if (colok) {
#pragma unroll
                for (int nt = 0; nt < kMaxTiles2; ++nt)
                    if (nt < ntile) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int row = nt * 16 + Mma<T>::row(lk, r);
                            if (row < len) {
                                const int nn = am + row;
                                X[(nn * nn + nn + m) * LDX + col] = c4[nt][r];
                            }
                        }
                    }
            }
        }
        __syncthreads();

        // stage 3: W_n = E_n Z_n (A[i][k] = E[i][k] = Et[k][i]), times e^{-i m phi}, accumulate
#pragma unroll
        for (int uu = 0; uu < MAXU; ++uu) {
            const int u = wave + 4 * uu;
            if (u < a.n_units3 && (a.units3[u] & 255) <= p) {
                const int code = a.units3[u];
                const int n = code & 255, mt = (code >> 8) & 255, ct = code >> 16;
                const int w = 2 * n + 1, base = n * n;
                const int roff = n * (2 * n - 1) * (2 * n + 1) / 3;
                const T* Etn = Etg + roff;      // global: transposed table, coalesced
                const T* Esn = Es + roff;       // LDS: row-major block, read transposed
                const int mo = mt * 16 + li;
                const int col = ct * 16 + li;
                const bool colok = col < NCOL;
                V4 c4 = V4{0, 0, 0, 0};
                for (int k0 = 0; k0 < w; k0 += 4 * KC) {
                    T av[KC], bv[KC];
#pragma unroll
                    for (int q = 0; q < KC; ++q) {
                        const int kin = k0 + 4 * q + lk;
                        av[q] = (mo < w && kin < w) ? (OPS ? Esn[mo * w + kin] : Etn[kin * w + mo]) : (T)0;
                        bv[q] = (kin < w && colok) ? X[(base + kin) * LDX + col] : (T)0;
                    }
#pragma unroll
                    for (int q = 0; q < KC; ++q)
                        if (k0 + 4 * q < w) c4 = Mma<T>::mma(av[q], bv[q], c4);
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
        if (u < a.n_units3 && (a.units3[u] & 255) <= p) {
            const int code = a.units3[u];
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
                        if (row < 2 * n + 1) {
                            double* x = dst + 2 * (n * n + row) + ri;
                            if (item.w)
                                atomicAdd(x, (double)acc[uu][r]);
                            else
                                *x += (double)acc[uu][r];
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

template <typename T, int TB, int MAXU, bool OPS>
void launch_one(const M2LBlockArgs* a, size_t shm, hipStream_t s) {
    if (!a) {   // prepare: allow the dynamic LDS this instantiation may use
        int dev = 0, lds = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) lds = 65536;
        hipError_t e = hipFuncSetAttribute((const void*)m2l_block_kernel<T, TB, MAXU, OPS>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) fail(-3, std::string("hipFuncSetAttribute(m2l) failed: ") + hipGetErrorString(e));
        return;
    }
    hipLaunchKernelGGL((m2l_block_kernel<T, TB, MAXU, OPS>), dim3(a->n_items), dim3(256), shm, s, *a);
}

template <typename T, int TB, bool OPS>
void launch_tb(const M2LBlockArgs* a, size_t shm, int maxu, hipStream_t s) {
    if (maxu <= 8)
        launch_one<T, TB, 8, OPS>(a, shm, s);
    else if (maxu <= 12)
        launch_one<T, TB, 12, OPS>(a, shm, s);
    else if (maxu <= 16)
        launch_one<T, TB, 16, OPS>(a, shm, s);
    else if (maxu <= 24)
        launch_one<T, TB, 24, OPS>(a, shm, s);
    else if (maxu <= 32)
        launch_one<T, TB, 32, OPS>(a, shm, s);
    else
        fail(-9, "internal: too many M2L units per wave");
}

template <typename T, bool OPS>
void launch_t(const M2LBlockArgs* a, size_t shm, int tb, int maxu, hipStream_t s) {
    if (tb == 16)
        launch_tb<T, 16, OPS>(a, shm, maxu, s);
    else if (tb == 8)
        launch_tb<T, 8, OPS>(a, shm, maxu, s);
    else
        launch_tb<T, 4, OPS>(a, shm, maxu, s);
}

void dispatch(const M2LBlockArgs* a, int p, int n_units3, int tb, int fp32, int ops, hipStream_t s) {
    const int maxu = (n_units3 + 3) / 4;
    const size_t shm = m2l_block_lds_bytes(p, tb, fp32 ? 4 : 8, ops);
    if (fp32) {
        if (ops)
            launch_t<float, true>(a, shm, tb, maxu, s);
        else
            launch_t<float, false>(a, shm, tb, maxu, s);
    } else {
        if (ops)
            launch_t<double, true>(a, shm, tb, maxu, s);
        else
            launch_t<double, false>(a, shm, tb, maxu, s);
    }
}

}  // namespace

void prepare_m2l_block(int p, int n_units3, int tb, int fp32, int ops) {
    dispatch(nullptr, p, n_units3, tb, fp32, ops, nullptr);
}

void launch_m2l_block(const M2LBlockArgs& a, int tb, int fp32, int ops, hipStream_t s) {
    if (a.n_items <= 0) return;
    dispatch(&a, a.p_tab, a.n_units3, tb, fp32, ops, s);
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
