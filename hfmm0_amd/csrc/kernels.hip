// HIP kernels of the Helmholtz FMM for gfx950 (CDNA4, wave64).
//
// Expansion conventions (shared with tables.cpp): box radius R = sqrt(3) h;
//   Mh_n^m = sum_j q_j jhat_n(k rho_j)(rho_j/R)^n conj(Y_n^m(y_j))
//   phi(x) = sum Lh_n^m jhat_n(k r)(r/R)^n Y_n^m(x)
// with jhat_n(x) = j_n(x)(2n+1)!!/x^n, orthonormal Y_n^m (Condon-Shortley phase) and
// coefficient index n*n + n + m.  Expansions are complex FP64, one contiguous
// (p+1)^2 vector per box.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "common.h"
#include "kernels.h"
#include "p2p_math.h"

namespace hfmm {

namespace {

constexpr int kXlateThreads = 256;
constexpr int kAccPerThread = ((kMaxP + 1) * (kMaxP + 1) + kXlateThreads - 1) / kXlateThreads;

__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cmadd(double2 acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
    return acc;
}
__device__ inline double2 conjd(double2 a) { return make_double2(a.x, -a.y); }

__device__ inline void nm_decode(int c, int& n, int& m) {
    n = (int)sqrtf((float)c + 0.5f);
    if ((n + 1) * (n + 1) <= c) ++n;
    if (n * n > c) --n;
    m = c - n * n - n;
}

// Legendre recurrence coefficients (fully normalised, Condon-Shortley phase):
//   P_m^m = -d_m sin(theta) P_{m-1}^{m-1},  d_m = sqrt((2m+1)/(2m))
//   P_{m+1}^m = e_m cos(theta) P_m^m,       e_m = sqrt(2m+3)
//   P_n^m = a_nm (cos(theta) P_{n-1}^m - b_nm P_{n-2}^m)
// stored order-major, ab[m][n] = (a_nm, b_nm), so that the coefficients of one order are
// contiguous; de[m] = (d_m, e_m).  Also the downward-recurrence factors of jhat,
// miller[n] = 1 / ((2n+1)(2n+3)).  The P2M and L2P kernels copy the part their degree
// needs into LDS once per workgroup: in the recurrence loops a constant-memory read at a
// lane-varying index is a vector memory load, and even a uniform one waits on the scalar
// cache, which this table (31 KB) does not fit.
constexpr int kMillerN = 512;
constexpr int kMillerLds = 128;   // factors staged in LDS (n < 128: k R below about 70 at p = 11)
struct LegCoef {
    double2 ab[kMaxP + 1][kMaxP + 1];
    double2 de[kMaxP + 1];
    double cm[kMaxP + 1];     // c_m = prod_{i <= m} (-d_i) / sqrt(4 pi): P_m^m = c_m sin^m(theta)
    double miller[kMillerN];
};
__constant__ LegCoef c_leg;

__device__ __forceinline__ double miller_factor(const double* mil, int nmil, int n) {
    return n < nmil ? mil[n] : 1.0 / ((2.0 * n + 1.0) * (2.0 * n + 3.0));
}

// jhat_n(x) for n = 0..p by Miller's downward recurrence
//   jhat_{n-1} = jhat_n - x^2 jhat_{n+1} / ((2n+1)(2n+3)),
// normalised with jhat_0 = sin(x)/x (or jhat_1 = 3(sin x - x cos x)/x^3 when that is
// the larger one).  out[n * stride].  The recurrence starts at N = p + 1.5 xmax + 12 for
// an upper bound xmax >= x common to the caller's workgroup (k R of the box), so every lane
// walks the same n and the factor table mil[0..nmil) is read at a uniform index.  Against
// SciPy's spherical Bessel functions (p = 2..40, x = 1e-3..150) a start of p + 1.5 x + 8
// already reaches the rounding floor (7e-14 relative to the neighbouring orders'
// magnitude); + 6 gives 2e-13, + 4 gives 9e-11.
__device__ void jhat_dev(double x, int p, double* out, int stride, double xmax, const double* mil, int nmil) {
    x = fabs(x);
    if (x == 0.0) {
        for (int n = 0; n <= p; ++n) out[n * stride] = 1.0;
        return;
    }
    const int N = p + (int)(1.5 * fmax(xmax, x)) + 12;
    double a_next = 0.0, a = 1.0;
    const double x2 = x * x;
    for (int n = N; n >= 1; --n) {
        const double am1 = a - x2 * a_next * miller_factor(mil, nmil, n);
        a_next = a;
        a = am1;
        if (n - 1 <= p) out[(n - 1) * stride] = a;
        if (fabs(a) > 1e250) {
            a *= 1e-250;
            a_next *= 1e-250;
            for (int j = n - 1; j <= p; ++j)
                if (j >= 0) out[j * stride] *= 1e-250;
        }
    }
    double s, c;
    sincos(x, &s, &c);
    const double e0 = s / x;
    double scale;
    if (x < 1.0) {
        scale = e0 / out[0];
    } else {
        const double e1 = 3.0 * (s - x * c) / (x * x * x);
        scale = fabs(e0) >= fabs(e1) ? e0 / out[0] : e1 / out[stride];
    }
    for (int n = 0; n <= p; ++n) out[n * stride] *= scale;
}

// LDS copy of the recurrence coefficients of degree p: ab in order-major triangle layout
// (the entries of order m, n = m..p, from leg_col(p + 1, m)), de and cm (p+1 each), and
// the first kMillerLds Miller factors; rounded up to 32 bytes, so that the double2 arrays
// the kernels place after it stay 16-byte aligned (3 (p+1) is odd for even p: misaligned
// 16-byte LDS accesses made L2P 3.8x slower at p = 12 than at p = 11)
__host__ __device__ constexpr int leg_col(int P1, int m) { return m * P1 - m * (m - 1) / 2; }
__host__ __device__ constexpr int leg_lds_doubles(int p) {
    return ((p + 1) * (p + 2) + 3 * (p + 1) + kMillerLds + 3) / 4 * 4;
}

struct LegLds {
    const double2* ab;    // ab[leg_col(p + 1, m) + n - m] = (a_nm, b_nm)
    const double2* de;    // (d_m, e_m)
    const double* cm;     // c_m
    const double* mil;    // Miller factors, n < kMillerLds
};

__device__ LegLds stage_legendre(double* dst, int p, int tid, int nthreads) {
    const int P1 = p + 1;
    double2* ab = (double2*)dst;
    double2* de = ab + P1 * (P1 + 1) / 2;
    double* cm = (double*)(de + P1);
    double* mil = cm + P1;
    for (int t = tid; t < P1 * P1; t += nthreads) {
        const int m = t / P1, n = t - m * P1;
        if (n >= m) ab[leg_col(P1, m) + n - m] = c_leg.ab[m][n];
    }
    for (int t = tid; t < P1; t += nthreads) {
        de[t] = c_leg.de[t];
        cm[t] = c_leg.cm[t];
    }
    for (int t = tid; t < kMillerLds; t += nthreads) mil[t] = c_leg.miller[t];
    return LegLds{ab, de, cm, mil};
}

// jhat_n(x), n = 0..p, up to one factor: out[n * stride] = jhat_n(x) / scale, with the
// returned scale (x >= 0).  Miller's downward recurrence
//   a_{n-1} = a_n - x^2 a_{n+1} / ((2n+1)(2n+3))
// from a_N = 1, a_{N+1} = 0 with a start N common to the workgroup (uniform loop: the
// caller gives N = p + 1.5 xmax + 13 for xmax >= x, as jhat_dev), normalised by
// jhat_0 = sin(x)/x or jhat_1 = 3(sin x - x cos x)/x^3 (the larger).  The factors come from
// mil[] (uniform LDS reads, four per wait) below kMillerLds.  x^2 f_n a_{n+1} depends on
// the previous step only through a_{n+1}, so one add per step is on the dependency chain.
// Values are brought back by 2^-830 when one passes 2^830 (where nothing is stored yet,
// tested every 4 steps: a step grows them by a factor of at most 1 + x^2 / 15).
__device__ __forceinline__ double miller_step_factor(const double* mil, int n) {
    return n < kMillerLds ? mil[n] : 1.0 / ((2.0 * n + 1.0) * (2.0 * n + 3.0));
}

__device__ double jhat_unscaled(double x, int p, double* out, int stride, int N, const double* mil) {
    if (x == 0.0) {
        for (int n = 0; n <= p; ++n) out[n * stride] = 1.0;
        return 1.0;
    }
    const double x2 = x * x;
    double a_next = 0.0, a = 1.0;
    int n = N;
    // steps n = N .. p + 2: nothing is stored
    for (; n - 3 >= p + 2; n -= 4) {
        const double f0 = miller_step_factor(mil, n), f1 = miller_step_factor(mil, n - 1);
        const double f2 = miller_step_factor(mil, n - 2), f3 = miller_step_factor(mil, n - 3);
        double t = a - x2 * f0 * a_next;
        a_next = a;
        a = t;
        t = a - x2 * f1 * a_next;
        a_next = a;
        a = t;
        t = a - x2 * f2 * a_next;
        a_next = a;
        a = t;
        t = a - x2 * f3 * a_next;
        a_next = a;
        a = t;
        if (fabs(a) > 0x1p830) {
            a *= 0x1p-830;
            a_next *= 0x1p-830;
        }
    }
    for (; n >= p + 2; --n) {
        const double t = a - x2 * miller_step_factor(mil, n) * a_next;
        a_next = a;
        a = t;
    }
    // steps n = p + 1 .. 1 store a_{n-1}
    for (; n >= 1; --n) {
        const double t = a - x2 * miller_step_factor(mil, n) * a_next;
        a_next = a;
        a = t;
        out[(n - 1) * stride] = a;
        if (fabs(a) > 0x1p830) {
            a *= 0x1p-830;
            a_next *= 0x1p-830;
            for (int j = n - 1; j <= p; ++j) out[j * stride] *= 0x1p-830;
        }
    }
    // a = a_0, a_next = a_1
    double sn, cs;
    sincos(x, &sn, &cs);
    if (x < 1.0) return sn / (x * a);
    const double e0 = sn / x, e1 = 3.0 * (sn - x * cs) / (x * x * x);
    return fabs(e0) >= fabs(e1) ? e0 / a : e1 / a_next;
}

// ---------------------------------------------------------------------------------
// small utility kernels
// ---------------------------------------------------------------------------------
__global__ void fill_zero_kernel(double2* buf, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) buf[i] = make_double2(0.0, 0.0);
}

__global__ void gather_charges_kernel(const void* q, int q_type, const int* perm, int64_t n, int conj,
                                      double2* qs) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = perm[i];
    double2 v;
    if (q_type == 0) {
        v = make_double2(((const double*)q)[o], 0.0);
    } else if (q_type == 1) {
        v = ((const double2*)q)[o];
    } else if (q_type == 2) {
        v = make_double2((double)((const float*)q)[o], 0.0);
    } else {
        const float2 f = ((const float2*)q)[o];
        v = make_double2((double)f.x, (double)f.y);
    }
    if (conj) v.y = -v.y;
    qs[i] = v;
}

// ---------------------------------------------------------------------------------
// Persistent leaf loops (P2M, L2P).  The grid is a multiple of 8 resident workgroups;
// workgroup i runs on XCD x = i % 8 and takes, in turn, the leaves l, l + g, l + 2g, ...
// (l = i / 8, g = grid / 8) of XCD x's contiguous range of n / 8 (+1 for x < n % 8)
// leaves, so that an XCD's L2 holds one region.  A leaf's descriptor is one 64-byte
// vector load by 16 lanes whose words are read back with v_readlane; the next leaf's load
// is issued before the current leaf is processed.
// ---------------------------------------------------------------------------------
struct LeafInfo {
    int box, level, pstart, count;
    int64_t exp_off;
    double cx, cy, cz;
};

__device__ __forceinline__ int leaf_desc_word(const LeafDesc* d, int it) {
    return ((const int*)(d + it))[threadIdx.x & 15];
}

__device__ __forceinline__ uint64_t readlane64(int w, int lo) {
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(w, lo + 1) << 32) |
           (uint32_t)__builtin_amdgcn_readlane(w, lo);
}

__device__ __forceinline__ LeafInfo leaf_desc_unpack(int w) {
    LeafInfo f;
    f.box = __builtin_amdgcn_readlane(w, 0);
    f.level = __builtin_amdgcn_readlane(w, 1);
    f.pstart = __builtin_amdgcn_readlane(w, 2);
    f.count = __builtin_amdgcn_readlane(w, 3);
    f.exp_off = (int64_t)readlane64(w, 4);
    f.cx = __builtin_bit_cast(double, readlane64(w, 8));
    f.cy = __builtin_bit_cast(double, readlane64(w, 10));
    f.cz = __builtin_bit_cast(double, readlane64(w, 12));
    return f;
}

struct LeafRange {
    int it, end, step;
};

__device__ __forceinline__ LeafRange leaf_range(int n) {
    const int g = gridDim.x / kXcds, x = blockIdx.x % kXcds, l = blockIdx.x / kXcds;
    const int q = n / kXcds, r = n % kXcds;
    const int beg = x * q + min(x, r);
    return LeafRange{beg + l, beg + q + (x < r ? 1 : 0), g};
}

// ---------------------------------------------------------------------------------
// P2M and L2P in solid-harmonic form.  With (x, y, z) a point relative to the box centre,
// rho its distance and R the expansion radius,
//   (rho/R)^n P_n^m(cos theta) e^{i m alpha} = Z_m T_n^m,
//   Z_m = c_m ((x + i y)/R)^m,   T_m^m = 1,  T_{m+1}^m = e_m z/R,
//   T_n^m = a_nm (z/R T_{n-1}^m - b_nm (rho/R)^2 T_{n-2}^m),
// the Legendre recurrence multiplied by (rho/R)^n: no angles and no divisions per point.
// jhat_n(k rho) comes from jhat_unscaled, whose factor joins the charge (P2M) or the sum
// (L2P).
// ---------------------------------------------------------------------------------

// ---------------------------------------------------------------------------------
// P2M: a workgroup per leaf at a time; points processed in chunks of up to 64 in LDS.
//  0. the recurrence coefficients of the largest leaf degree are copied into LDS once;
//  1. one thread per point: jhat_n(k rho) up to a factor, which joins the charge
//     (q' = q scale), z/R, (rho/R)^2, and Z_m for m = 0..p;
//  2. one thread per (order m, block of up to 4 degrees n) and point group g (points
//     j = g mod G): per point T_n^m from T_m^m = 1, and for each (n, m) of the block, with
//     w_j = q'_j jhat_n T_n^m, the sums U = sum_j w_j Re Z_m and V = sum_j w_j Im Z_m,
//     which give both orders: Mh_n^m = U - iV and Mh_n^-m = (-1)^m (U + iV), since
//     conj(Y_n^-m) = (-1)^m Y_n^m.  The G groups are added at the end of the leaf.
// ---------------------------------------------------------------------------------
constexpr int kP2MBlockN = 4;     // degrees per thread block of an order

__host__ __device__ constexpr int p2m_blocks(int p) {
    int nb = 0;
    for (int m = 0; m <= p; ++m) nb += (p + 1 - m + kP2MBlockN - 1) / kP2MBlockN;
    return nb;
}
// threads of a workgroup: 128 while every block of the largest leaf degree gets a thread,
// else 256 (p2m_blocks(40) = 231).  128 keeps 7 leaves in flight per CU against 4 for 256,
// which the latency of the per-point recurrences needs (P2M at N = 1M, k = 40: 0.32 against
// 0.47 ms); 64 leaves each thread three times the points and measured 0.45 ms
__host__ __device__ constexpr int p2m_threads(int pmax) { return p2m_blocks(pmax) <= 128 ? 128 : 256; }
// point groups: blocks x groups fill the workgroup of NT threads, at most 16
__host__ __device__ constexpr int p2m_groups(int p, int NT) {
    return NT / p2m_blocks(p) < 16 ? (NT / p2m_blocks(p) > 0 ? NT / p2m_blocks(p) : 1) : 16;
}

// Miller start for a leaf (uniform): p + 1.5 k R + 13
__device__ __forceinline__ int miller_start(int p, double kR) { return p + (int)(1.5 * fabs(kR)) + 13; }

// held to 4 waves per SIMD (128 VGPRs)
template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void p2m_kernel(
    const LeafDesc* __restrict__ leaves, int n_items, int chunk, int pmax, DevLevels lv, const double* X,
    const double* Y, const double* Z, const double2* q, double k, double2* M) {
    extern __shared__ double smem[];
    const int tid = threadIdx.x;
    const int PM1 = pmax + 1;
    const LegLds lg = stage_legendre(smem, pmax, tid, NT);
    double* region = smem + leg_lds_doubles(pmax);
    LeafRange lr = leaf_range(n_items);
    int w = lr.it < lr.end ? leaf_desc_word(leaves, lr.it) : 0;
    __syncthreads();
    for (; lr.it < lr.end; lr.it += lr.step) {
        const LeafInfo f = leaf_desc_unpack(w);
        if (lr.it + lr.step < lr.end) w = leaf_desc_word(leaves, lr.it + lr.step);
        const int p = lv.p[f.level];
        const double R = lv.R[f.level];
        const double invR = 1.0 / R;
        const int N = miller_start(p, k * R);
        const int P1 = p + 1;
        double2* s_z = (double2*)region;              // chunk x P1: Z_m
        double2* s_q = s_z + chunk * P1;              // chunk: q' = q scale
        double2* s_g = s_q + chunk;                   // chunk: (z/R, (rho/R)^2)
        double* s_j = (double*)(s_g + chunk);         // chunk x P1: jhat_n / scale
        // this thread's block: order m, degrees n0..n1, point group g (g >= G: idle)
        const int NB = p2m_blocks(p), G = p2m_groups(p, NT);
        const int bi = tid % NB, g = tid / NB;
        int m = 0, n0 = 0;
        for (int r = bi;; ++m) {
            const int nbm = (p + 1 - m + kP2MBlockN - 1) / kP2MBlockN;
            if (r < nbm) {
                n0 = m + r * kP2MBlockN;
                break;
            }
            r -= nbm;
        }
        const int n1 = min(p, n0 + kP2MBlockN - 1);
        const bool active = g < G;
        const double2* abm = lg.ab + leg_col(PM1, m) - m;   // (a_nm, b_nm) at abm[n], n >= m + 2
        double2 U[kP2MBlockN], V[kP2MBlockN];
#pragma unroll
        for (int s = 0; s < kP2MBlockN; ++s) U[s] = V[s] = make_double2(0.0, 0.0);

        for (int j0 = 0; j0 < f.count; j0 += chunk) {
            const int nch = min(chunk, f.count - j0);
            for (int j = tid; j < nch; j += NT) {
                const int i = f.pstart + j0 + j;
                const double dx = X[i] - f.cx, dy = Y[i] - f.cy, dz = Z[i] - f.cz;
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double scale = jhat_unscaled(k * sqrt(r2), p, s_j + j * P1, 1, N, lg.mil);
                const double2 qq = q[i];
                s_q[j] = make_double2(qq.x * scale, qq.y * scale);
                s_g[j] = make_double2(dz * invR, r2 * invR * invR);
                const double2 u = make_double2(dx * invR, dy * invR);
                double2 um = make_double2(1.0, 0.0);   // ((x + i y)/R)^m
                s_z[j * P1] = make_double2(lg.cm[0], 0.0);
                for (int mm = 1; mm <= p; ++mm) {
                    um = cmul(um, u);
                    const double c = lg.cm[mm];
                    s_z[j * P1 + mm] = make_double2(c * um.x, c * um.y);
                }
            }
            __syncthreads();
            if (active) {
                const double e_m = lg.de[m].y;
                // two points per pass: two independent recurrences and their LDS reads
                // interleave (a missing second point repeats the first with weight 0)
                for (int j = g; j < nch; j += 2 * G) {
                    const bool two = j + G < nch;
                    const int j2 = two ? j + G : j;
                    const double2 zA = s_z[j * P1 + m], zB = s_z[j2 * P1 + m];
                    const double2 qA = s_q[j], qB = two ? s_q[j2] : make_double2(0.0, 0.0);
                    const double2 gA = s_g[j], gB = s_g[j2];
                    const double* jA = s_j + j * P1;
                    const double* jB = s_j + j2 * P1;
                    // T_n^m at n = n0 (t1) and n0 - 1 (t2)
                    double a1 = 1.0, a2 = 0.0, b1 = 1.0, b2 = 0.0;
                    for (int n = m + 1; n <= n0; ++n) {
                        const double2 cf = n == m + 1 ? make_double2(e_m, 0.0) : abm[n];
                        const double ta = cf.x * (gA.x * a1 - cf.y * gA.y * a2);
                        const double tb = cf.x * (gB.x * b1 - cf.y * gB.y * b2);
                        a2 = a1;
                        a1 = ta;
                        b2 = b1;
                        b1 = tb;
                    }
#pragma unroll
                    for (int s = 0; s < kP2MBlockN; ++s) {
                        const int n = n0 + s;
                        if (n > n1) break;
                        if (s > 0) {
                            const double2 cf = n == m + 1 ? make_double2(e_m, 0.0) : abm[n];
                            const double ta = cf.x * (gA.x * a1 - cf.y * gA.y * a2);
                            const double tb = cf.x * (gB.x * b1 - cf.y * gB.y * b2);
                            a2 = a1;
                            a1 = ta;
                            b2 = b1;
                            b1 = tb;
                        }
                        const double hA = jA[n] * a1, hB = jB[n] * b1;
                        const double wrA = qA.x * hA, wiA = qA.y * hA, wrB = qB.x * hB, wiB = qB.y * hB;
                        U[s].x = fma(wrA, zA.x, fma(wrB, zB.x, U[s].x));
                        U[s].y = fma(wiA, zA.x, fma(wiB, zB.x, U[s].y));
                        V[s].x = fma(wrA, zA.y, fma(wrB, zB.y, V[s].x));
                        V[s].y = fma(wiA, zA.y, fma(wiB, zB.y, V[s].y));
                    }
                }
            }
            __syncthreads();
        }
        // the groups' partial sums meet in LDS (the chunk region is free now); group 0 writes
        if (G > 1) {
            double2* red = (double2*)region;   // (G - 1) x NB x kP2MBlockN x (U, V)
            if (g >= 1 && active)
#pragma unroll
                for (int s = 0; s < kP2MBlockN; ++s) {
                    const int o = 2 * (((g - 1) * NB + bi) * kP2MBlockN + s);
                    red[o] = U[s];
                    red[o + 1] = V[s];
                }
            __syncthreads();
            if (g == 0)
                for (int h = 1; h < G; ++h)
#pragma unroll
                    for (int s = 0; s < kP2MBlockN; ++s) {
                        const int o = 2 * (((h - 1) * NB + bi) * kP2MBlockN + s);
                        const double2 u = red[o], v = red[o + 1];
                        U[s].x += u.x;
                        U[s].y += u.y;
                        V[s].x += v.x;
                        V[s].y += v.y;
                    }
            __syncthreads();   // the next leaf's chunk overwrites red
        }
        if (g == 0) {
            double2* out = M + f.exp_off;
            const double sg = (m & 1) ? -1.0 : 1.0;
#pragma unroll
            for (int s = 0; s < kP2MBlockN; ++s) {
                const int n = n0 + s;
                if (n > n1) break;
                const double2 u = U[s], v = V[s];
                out[n * n + n + m] = make_double2(u.x + v.y, u.y - v.x);
                if (m > 0) out[n * n + n - m] = make_double2(sg * (u.x - v.y), sg * (u.y + v.x));
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// L2P: a workgroup (one wave) per leaf at a time.  The recurrence coefficients (once) and
// the leaf's local expansion, recombined per (n, m >= 0) as
//   A_nm = Lh_n^m + (-1)^m Lh_n^-m,  B_nm = Lh_n^m - (-1)^m Lh_n^-m  (B_n0 = 0),
// are staged in LDS (Y_n^-m = (-1)^m conj(Y_n^m)).  A leaf of at most 16 (32) points
// gives each point 4 (2) lanes, which split the orders m between them (zig-zag, so that
// the parts get about equal numbers of (n, m) terms) and add their results by shuffles.
// For each of its orders m a lane sums S = sum_n j_n T_n^m A_nm and T = sum_n j_n T_n^m B_nm
// (j_n = jhat_n(k rho) / scale, kept in LDS as [n][lane]) and adds
// S Re Z_m + i T Im Z_m; the sum is multiplied by the scale at the end.
// ---------------------------------------------------------------------------------
__host__ __device__ constexpr int l2p_lds_doubles(int p) {
    return leg_lds_doubles(p) + 2 * (p + 1) * (p + 2) + 64 * (p + 1);
}

__global__ __launch_bounds__(64) void l2p_kernel(const LeafDesc* __restrict__ leaves, int n_items, int pmax,
                                                 DevLevels lv, const double* X, const double* Y, const double* Z,
                                                 const double2* Lexp, double k, double2* far) {
    extern __shared__ double smem[];
    const int lane = threadIdx.x;
    const int PM1 = pmax + 1;
    const LegLds lg = stage_legendre(smem, pmax, lane, 64);
    double2* s_AB = (double2*)(smem + leg_lds_doubles(pmax));    // (A, B) per (m, n), order-major triangle
    double* s_rad = (double*)(s_AB + PM1 * (PM1 + 1));           // [n][lane]
    double* rad = s_rad + lane;
    LeafRange lr = leaf_range(n_items);
    int w = lr.it < lr.end ? leaf_desc_word(leaves, lr.it) : 0;
    for (; lr.it < lr.end; lr.it += lr.step) {
        const LeafInfo f = leaf_desc_unpack(w);
        if (lr.it + lr.step < lr.end) w = leaf_desc_word(leaves, lr.it + lr.step);
        const int p = lv.p[f.level];
        const double R = lv.R[f.level];
        const double invR = 1.0 / R;
        const int N = miller_start(p, k * R);
        const int P1 = p + 1;
        const double2* src = Lexp + f.exp_off;
        __syncthreads();   // the previous leaf's reads of s_AB are done (and the staging at the start)
        for (int t = lane; t < P1 * P1; t += 64) {
            const int m = t / P1, n = t - m * P1;
            if (n < m) continue;
            const double2 lp = src[n * n + n + m];
            double2 A = lp, B = make_double2(0.0, 0.0);
            if (m > 0) {
                const double2 lm = src[n * n + n - m];
                const double sg = (m & 1) ? -1.0 : 1.0;
                A = make_double2(lp.x + sg * lm.x, lp.y + sg * lm.y);
                B = make_double2(lp.x - sg * lm.x, lp.y - sg * lm.y);
            }
            const int c = leg_col(P1, m) + n - m;
            s_AB[2 * c] = A;
            s_AB[2 * c + 1] = B;
        }
        __syncthreads();
        const int cnt = f.count;
        const int S = cnt <= 16 ? 4 : (cnt <= 32 ? 2 : 1);   // lanes per point
        const int W = 64 / S;                                // points per pass
        const int part = lane / W;
        for (int i0 = 0; i0 < cnt; i0 += W) {
            const int il = i0 + lane % W;
            const bool valid = il < cnt;
            const int i = f.pstart + (valid ? il : cnt - 1);
            const double dx = X[i] - f.cx, dy = Y[i] - f.cy, dz = Z[i] - f.cz;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double scale = jhat_unscaled(k * sqrt(r2), p, rad, 64, N, lg.mil);
            const double zr = dz * invR, rr2 = r2 * invR * invR;
            const double2 u = make_double2(dx * invR, dy * invR);
            double2 phi = make_double2(0.0, 0.0);
            double2 um = make_double2(1.0, 0.0);   // ((x + i y)/R)^m
            for (int m = 0; m <= p; ++m) {
                if (m > 0) um = cmul(um, u);
                const int mr = m % S;
                if (((m / S) & 1 ? S - 1 - mr : mr) != part) continue;
                const double2* abm = lg.ab + leg_col(PM1, m) - m;     // (a_nm, b_nm) at abm[n]
                const double2* ABm = s_AB + 2 * (leg_col(P1, m) - m);  // (A_nm, B_nm) at ABm[2n], ABm[2n + 1]
                double h = rad[64 * m];
                double2 Sv = make_double2(h * ABm[2 * m].x, h * ABm[2 * m].y);
                double2 Tv = make_double2(h * ABm[2 * m + 1].x, h * ABm[2 * m + 1].y);
                if (m < p) {
                    double t2 = 1.0, t1 = lg.de[m].y * zr;
                    h = rad[64 * (m + 1)] * t1;
                    Sv.x = fma(h, ABm[2 * m + 2].x, Sv.x);
                    Sv.y = fma(h, ABm[2 * m + 2].y, Sv.y);
                    Tv.x = fma(h, ABm[2 * m + 3].x, Tv.x);
                    Tv.y = fma(h, ABm[2 * m + 3].y, Tv.y);
#pragma unroll 2
                    for (int n = m + 2; n <= p; ++n) {
                        const double2 cf = abm[n];
                        const double tn = cf.x * (zr * t1 - cf.y * rr2 * t2);
                        t2 = t1;
                        t1 = tn;
                        h = rad[64 * n] * tn;
                        const double2 A = ABm[2 * n], B = ABm[2 * n + 1];
                        Sv.x = fma(h, A.x, Sv.x);
                        Sv.y = fma(h, A.y, Sv.y);
                        Tv.x = fma(h, B.x, Tv.x);
                        Tv.y = fma(h, B.y, Tv.y);
                    }
                }
                const double c = lg.cm[m];
                const double zx = c * um.x, zy = c * um.y;
                phi.x += Sv.x * zx - Tv.y * zy;
                phi.y += Sv.y * zx + Tv.x * zy;
            }
            // the parts of a point are W lanes apart
            for (int off = W; off < 64; off <<= 1) {
                phi.x += __shfl_xor(phi.x, off);
                phi.y += __shfl_xor(phi.y, off);
            }
            if (valid && part == 0) far[f.pstart + il] = make_double2(phi.x * scale, phi.y * scale);
        }
    }
}

// ---------------------------------------------------------------------------------
// L2P with the field.  The gradient of the scaled regular basis
// B_nm = jhat_n(kr) (r/R)^n Y_n^m stays in the basis, one degree up or down:
//   d/dz B_nm            = az(n,m) B_{n-1,m}   + gz(n,m) B_{n+1,m}
//   (d/dx + i d/dy) B_nm = ap(n,m) B_{n-1,m+1} + gp(n,m) B_{n+1,m+1}
//   (d/dx - i d/dy) B_nm = am(n,m) B_{n-1,m-1} + gm(n,m) B_{n+1,m-1}
// (the gradient formulas of j_n Y_n^m rescaled by (2n+1)!!/(kR)^n; grad_lo / grad_hi
// give the coefficients).  A workgroup (one wave) first forms the three derivative
// expansions (degree p+1) of the leaf's local expansion in LDS, then one thread per
// point evaluates phi and the derivatives; E = -grad phi.
// ---------------------------------------------------------------------------------
// coefficient of B_{n-1,.} in derivative `which` (0: d/dz, 1: d/dx + i d/dy, 2: d/dx - i d/dy) of B_nm
__device__ inline double grad_lo(int which, int n, int m, double R) {
    if (n < 1) return 0.0;
    const double num = which == 0 ? (double)(n + m) * (n - m)
                                  : (which == 1 ? (double)(n - m - 1) * (n - m) : (double)(n + m - 1) * (n + m));
    const double v = (2.0 * n + 1.0) / R * sqrt(num / ((2.0 * n - 1.0) * (2.0 * n + 1.0)));
    return which == 2 ? -v : v;
}
// coefficient of B_{n+1,.}
__device__ inline double grad_hi(int which, int n, int m, double k, double R) {
    const double num = which == 0 ? (double)(n + 1 + m) * (n + 1 - m)
                                  : (which == 1 ? (double)(n + m + 1) * (n + m + 2) : (double)(n - m + 1) * (n - m + 2));
    const double v = k * k * R / (2.0 * n + 3.0) * sqrt(num / ((2.0 * n + 1.0) * (2.0 * n + 3.0)));
    return which == 1 ? v : -v;
}

__global__ __launch_bounds__(64) void l2p_field_kernel(const int* leaves, DevBoxes bx, DevLevels lv,
                                                       const double* X, const double* Y, const double* Z,
                                                       const double2* Lexp, double k, double2* far,
                                                       double2* far_e) {
    extern __shared__ double smem[];
    const int b = leaves[blockIdx.x];
    const int L = bx.level[b];
    const int p = lv.p[L];
    const double R = lv.R[L];
    const int P1 = p + 1, P2 = p + 2, nc = P1 * P1, nd = P2 * P2;
    double2* s_L = (double2*)smem;
    double2* s_D = s_L + nc;                 // 3 sets of nd: d/dz, d/dx + i d/dy, d/dx - i d/dy
    double* s_rad = (double*)(s_D + 3 * nd);
    const double2* src = Lexp + bx.exp_off[b];
    for (int c = threadIdx.x; c < nc; c += blockDim.x) s_L[c] = src[c];
    __syncthreads();
    for (int w = threadIdx.x; w < 3 * nd; w += blockDim.x) {
        const int which = w / nd, c = w - which * nd;
        int n, m;
        nm_decode(c, n, m);
        const int ms = which == 0 ? m : (which == 1 ? m - 1 : m + 1);   // order of the source terms
        double2 d = make_double2(0.0, 0.0);
        const int am = ms < 0 ? -ms : ms;
        if (n + 1 <= p && am <= n + 1) {
            const double f = grad_lo(which, n + 1, ms, R);
            const double2 l = s_L[(n + 1) * (n + 1) + (n + 1) + ms];
            d.x += f * l.x;
            d.y += f * l.y;
        }
        if (n >= 1 && am <= n - 1) {
            const double f = grad_hi(which, n - 1, ms, k, R);
            const double2 l = s_L[(n - 1) * (n - 1) + (n - 1) + ms];
            d.x += f * l.x;
            d.y += f * l.y;
        }
        s_D[w] = d;
    }
    __syncthreads();
    const double cx = bx.centre[3 * b], cy = bx.centre[3 * b + 1], cz = bx.centre[3 * b + 2];
    const int ps = bx.pstart[b], cnt = bx.count[b];
    double* rad = s_rad + threadIdx.x * P2;
    for (int i0 = 0; i0 < cnt; i0 += blockDim.x) {
        const int il = i0 + threadIdx.x;
        if (il >= cnt) break;
        const int i = ps + il;
        const double dx = X[i] - cx, dy = Y[i] - cy, dz = Z[i] - cz;
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        const double rxy = sqrt(dx * dx + dy * dy);
        const double ct = r > 0.0 ? dz / r : 1.0;
        const double st = r > 0.0 ? rxy / r : 0.0;
        const double2 eia = rxy > 0.0 ? make_double2(dx / rxy, dy / rxy) : make_double2(1.0, 0.0);
        jhat_dev(k * r, p + 1, rad, 1, k * R, c_leg.miller, kMillerN);
        const double rr = r / R;
        double pw = 1.0;
        for (int n = 0; n <= p + 1; ++n) {
            rad[n] *= pw;
            pw *= rr;
        }
        double2 acc[4] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0),
                          make_double2(0.0, 0.0)};   // phi, d/dz, d/dx + i d/dy, d/dx - i d/dy
        double pmm = 0.28209479177387814;
        double2 eim = make_double2(1.0, 0.0);
        for (int m = 0; m <= p + 1; ++m) {
            if (m > 0) {
                pmm *= -c_leg.de[m].x * st;
                eim = cmul(eim, eia);
            }
            const double sgn = (m & 1) ? -1.0 : 1.0;
            double p2 = 0.0, p1 = 0.0;
            for (int n = m; n <= p + 1; ++n) {
                double pn;
                if (n == m)
                    pn = pmm;
                else if (n == m + 1)
                    pn = c_leg.de[m].y * ct * pmm;
                else
                    pn = c_leg.ab[m][n].x * (ct * p1 - c_leg.ab[m][n].y * p2);
                p2 = p1;
                p1 = pn;
                const double g = rad[n] * pn;
                for (int set = 0; set < 4; ++set) {
                    if (set == 0 && n > p) continue;
                    const double2* C = set == 0 ? s_L : s_D + (set - 1) * nd;
                    double2 term = cmul(C[n * n + n + m], eim);
                    if (m > 0) {
                        const double2 t2 = cmul(C[n * n + n - m], conjd(eim));
                        term.x += sgn * t2.x;
                        term.y += sgn * t2.y;
                    }
                    acc[set].x = fma(g, term.x, acc[set].x);
                    acc[set].y = fma(g, term.y, acc[set].y);
                }
            }
        }
        far[i] = acc[0];
        // d/dx = (D+ + D-)/2, d/dy = (D+ - D-)/(2i), E = -grad phi
        const double2 sp = acc[2], sm = acc[3];
        far_e[3 * (int64_t)i + 0] = make_double2(-0.5 * (sp.x + sm.x), -0.5 * (sp.y + sm.y));
        far_e[3 * (int64_t)i + 1] = make_double2(-0.5 * (sp.y - sm.y), 0.5 * (sp.x - sm.x));
        far_e[3 * (int64_t)i + 2] = make_double2(-acc[1].x, -acc[1].y);
    }
}

// ---------------------------------------------------------------------------------
// Translation (M2M, M2L, L2L): rotate the source expansion so the translation axis
// is +z, apply the coaxial operator (block diagonal in m), rotate back, accumulate.
// One workgroup per target box; thread per coefficient; the three stages exchange
// data through LDS.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(kXlateThreads) void xlate_kernel(XlateArgs a) {
    extern __shared__ double2 sm2[];
    const int ps = a.ps, pt = a.pt;
    const int ncs = (ps + 1) * (ps + 1), nct = (pt + 1) * (pt + 1);
    const int nmx = max(ncs, nct);
    double2* sIn = sm2;          // stage-0 output, reused for the stage-2 output
    double2* sA = sIn + nmx;
    double2* sB = sIn;
    int* s_off = (int*)(sA + nmx);
    const int tps = a.tps, tpt = a.tpt;   // degrees the coaxial table was packed for
    const int mmax = min(tps, tpt);
    if (threadIdx.x == 0) {
        int o = 0;
        for (int m = 0; m <= mmax; ++m) {
            s_off[m] = o;
            o += (tpt + 1 - m) * (tps + 1 - m);
        }
    }
    const int t_ord = blockIdx.x / a.split, part = blockIdx.x % a.split;
    const int tbox = a.targets[t_ord];
    const int e_lo = a.ptr[t_ord], e_n = a.ptr[t_ord + 1] - e_lo;
    const int e_begin = e_lo + (int)((int64_t)e_n * part / a.split);
    const int e_end = e_lo + (int)((int64_t)e_n * (part + 1) / a.split);
    const int rsz = (a.pmax + 1) * (2 * a.pmax + 1) * (2 * a.pmax + 3) / 3;
    double2 acc[kAccPerThread];
#pragma unroll
    for (int t = 0; t < kAccPerThread; ++t) acc[t] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int e = e_begin; e < e_end; ++e) {
        const int s = a.src[e];
        const int d = a.dir[e];
        const DispInfo di = a.dinfo[d];
        const double2* ph = a.phase + (size_t)d * (a.pmax + 1);
        const double* rot = a.rot + (size_t)di.theta * rsz;
        const double2* cx = a.coax + (size_t)di.dist * a.coax_stride;
        const double2* src = a.srcbuf + a.exp_off[s];
        for (int c = threadIdx.x; c < ncs; c += blockDim.x) {
            int n, m;
            nm_decode(c, n, m);
            const double2 phm = m >= 0 ? ph[m] : conjd(ph[-m]);
            sIn[c] = cmul(src[c], phm);
        }
        __syncthreads();
        for (int c = threadIdx.x; c < ncs; c += blockDim.x) {
            int n, m;
            nm_decode(c, n, m);
            const int w = 2 * n + 1;
            const double* E = rot + n * (2 * n - 1) * (2 * n + 1) / 3;
            double2 v = make_double2(0.0, 0.0);
            const double2* in = sIn + n * n;
            for (int j = 0; j < w; ++j) {
                const double e2 = E[j * w + (m + n)];
                v.x = fma(e2, in[j].x, v.x);
                v.y = fma(e2, in[j].y, v.y);
            }
            sA[c] = v;
        }
        __syncthreads();
        for (int c = threadIdx.x; c < nct; c += blockDim.x) {
            int n, m;
            nm_decode(c, n, m);
            const int am = m < 0 ? -m : m;
            double2 v = make_double2(0.0, 0.0);
            if (am <= ps) {
                const double2* T = cx + s_off[am] + (n - am) * (tps + 1 - am);
                for (int n2 = am; n2 <= ps; ++n2) v = cmadd(v, T[n2 - am], sA[n2 * n2 + n2 + m]);
            }
            sB[c] = v;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kAccPerThread; ++t) {
            const int c = threadIdx.x + t * blockDim.x;
            if (c < nct) {
                int n, m;
                nm_decode(c, n, m);
                const int w = 2 * n + 1;
                const double* E = rot + n * (2 * n - 1) * (2 * n + 1) / 3 + (m + n) * w;
                const double2* in = sB + n * n;
                double2 v = make_double2(0.0, 0.0);
                for (int j = 0; j < w; ++j) {
                    v.x = fma(E[j], in[j].x, v.x);
                    v.y = fma(E[j], in[j].y, v.y);
                }
                const double2 phm = m >= 0 ? ph[m] : conjd(ph[-m]);
                const double2 r = cmul(v, conjd(phm));
                acc[t].x += r.x;
                acc[t].y += r.y;
            }
        }
        __syncthreads();
    }
    double2* dst = a.dstbuf + a.exp_off[tbox];
    if (a.split > 1) {   // parts of one target's sources: added atomically
        if (e_begin == e_end) return;
#pragma unroll
        for (int t = 0; t < kAccPerThread; ++t) {
            const int c = threadIdx.x + t * blockDim.x;
            if (c < nct) {
                atomicAdd(&dst[c].x, acc[t].x);
                atomicAdd(&dst[c].y, acc[t].y);
            }
        }
        return;
    }
#pragma unroll
    for (int t = 0; t < kAccPerThread; ++t) {
        const int c = threadIdx.x + t * blockDim.x;
        if (c < nct) {
            if (a.accumulate) {
                double2 v = dst[c];
                v.x += acc[t].x;
                v.y += acc[t].y;
                dst[c] = v;
            } else {
                dst[c] = acc[t];
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// P2P near field.
// ---------------------------------------------------------------------------------
using nearfield::radian_weight;
using nearfield::radian_weight_phase64;

template <typename T>
struct P2PMath;

template <>
struct P2PMath<float> {
    __device__ static inline void kernel(float r2, float k, float& wr, float& wi) { radian_weight(r2, k, wr, wi); }
};

template <>
struct P2PMath<double> {
    __device__ static inline void kernel(double r2, double k, double& wr, double& wi) {
        const double r = sqrt(r2);
        const double rinv = 1.0 / r;
        double s, c;
        sincos(k * r, &s, &c);
        wr = c * rinv;
        wi = s * rinv;
    }
};

// e^{ikr}/r and the field weight (1 - ikr) e^{ikr}/r^3, so that
// -grad_x e^{ik|x-y|}/|x-y| = (x - y) * field weight (HFMM.py:29-37 in complex form)
template <typename T>
__device__ inline void field_weights(T r2, T k, T& wr, T& wi, T& fr, T& fi) {
    P2PMath<T>::kernel(r2, k, wr, wi);
    const T rinv2 = (T)1 / r2;
    const T kr = k * sqrt(r2);
    fr = (wr + kr * wi) * rinv2;
    fi = (wi - kr * wr) * rinv2;
}

// Source staging of the near-field kernels: lane `lane` of a batch starting at `base`
// walks the leaf's list of point ranges (prefix sums rng_pref over [r0, r1)) to the range
// holding source v = base + lane; `ri` carries the range index from batch to batch (it
// only grows), so the walk is a step or two per batch.
__device__ inline int p2p_source_index(int v, int r1, const int* rng_start, const int* rng_pref, int& ri) {
    while (ri + 1 < r1 && rng_pref[ri + 1] <= v) ++ri;
    return rng_start[ri] + (v - rng_pref[ri]);
}

// Target chunk of a leaf: up to 64 R targets; a remainder that is not a power of two
// (and above 4 R) goes as its largest power of two first, so that few target slots of
// the R x L lanes are idle (each chunk stages the sources once more).
template <int R>
__device__ inline int p2p_chunk(int remaining) {
    if (remaining >= 64 * R) return 64 * R;
    if (remaining <= 4 * R) return remaining;
    int p = 1;
    while (2 * p <= remaining) p <<= 1;
    return p;
}

// Mixed-precision near field: one wave per target leaf.  Each lane holds R targets of the
// chunk (li, li + L, ..., with L lanes per group, L = next power of two >= chunk / R) and
// scans every G-th source (G = 64 / L groups) of each 64-point batch staged in LDS, so
// that every source read from LDS feeds R independent pair evaluations.  Coordinates are
// relative to the target leaf centre and scaled by a.scale = k / pi, in FP64: the
// coordinate differences and r^2 are FP64 (exact to rounding for close pairs; FP32 at the
// FP32-phase leaves flagged in a.coord32, p2p_mixed_body_c32), the
// weights e^{i pi u}/u (u = k r / pi) FP32 (halfturn_weight; PH: phase in FP64), the
// 64-source batch sums FP32, added into FP64 sums per target.  LOWK (k r below 1e-4 over
// every pair of the plan, k = 0 included): coordinates unscaled and the weights
// 1/r + i k, exact to FP32 rounding.  Pairs with r^2 at or below the leaf's coincidence
// floor contribute nothing.  The sums times a.wscale (k / pi; 1 for LOWK) are added to the
// far field and written in input order.  The workgroup handles leaf a.leaf0 + blockIdx.x.
// LDS of the near-field kernels: 64 staged sources
union P2PStage {
    struct {
        double2 xy[64];
        double2 zq[64];   // (z, charge as two FP32 bit-cast into one double)
    } mixed;
    struct {
        double x[64], y[64], z[64], qr[64], qi[64];
    } fp64;
};

// The FP32-phase mixed near field of a leaf whose coordinate differences and r^2 may be
// FP32 (a.coord32): the same schedule as p2p_mixed_body, with the scaled coordinates
// relative to the leaf centre rounded to FP32 once (targets in registers, sources staged
// in LDS as (x, y, z) and the charge).
template <int R>
__device__ __forceinline__ void p2p_mixed_body_c32(const P2PArgs& a, const int leaf, P2PStage& st) {
    float4* sxyz = (float4*)st.mixed.xy;
    float2* sq = (float2*)st.mixed.zq;
    const int b = a.leaf_box[leaf];
    const int tstart = a.bx.pstart[b], tcnt = a.bx.count[b];
    const double sc = a.scale;
    const double cx = a.bx.centre[3 * b], cy = a.bx.centre[3 * b + 1], cz = a.bx.centre[3 * b + 2];
    const int r0 = a.rng_ptr[leaf], r1 = a.rng_ptr[leaf + 1];
    const int S = a.leaf_src_total[leaf];
    // the floor in scaled coordinates, at least 2^-120 (an FP32 normal number)
    const float fl2 = (float)fmax(a.floor2[leaf] * sc * sc, 0x1p-120);
    const int lane = threadIdx.x;
    for (int t0 = 0, tc = 0; t0 < tcnt; t0 += tc) {
        tc = p2p_chunk<R>(tcnt - t0);
        const int need = (tc + R - 1) / R;
        int L = 1;
        while (L < need) L <<= 1;
        const int G = 64 / L;
        const int li = lane & (L - 1), g = lane / L;
        float xt[R], yt[R], zt[R];
        double sr[R], si[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int it = tstart + t0 + min(li + r * L, tc - 1);
            xt[r] = (float)((a.X[it] - cx) * sc);
            yt[r] = (float)((a.Y[it] - cy) * sc);
            zt[r] = (float)((a.Z[it] - cz) * sc);
            sr[r] = 0.0;
            si[r] = 0.0;
        }
        int ri = r0;
        for (int base = 0; base < S; base += 64) {
            const int v = base + lane;
            if (v < S) {
                const int idx = p2p_source_index(v, r1, a.rng_start, a.rng_pref, ri);
                const double2 qq = a.q[idx];
                sxyz[lane] = make_float4((float)((a.X[idx] - cx) * sc), (float)((a.Y[idx] - cy) * sc),
                                         (float)((a.Z[idx] - cz) * sc), 0.0f);
                sq[lane] = make_float2((float)qq.x, (float)qq.y);
            }
            __syncthreads();
            const int nb = min(64, S - base);
            float ar[R], ai[R];
#pragma unroll
            for (int r = 0; r < R; ++r) ar[r] = ai[r] = 0.0f;
            for (int j = g; j < nb; j += G) {
                const float4 p = sxyz[j];
                const float2 qj = sq[j];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const float dx = xt[r] - p.x, dy = yt[r] - p.y, dz = zt[r] - p.z;
                    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    float c, s, rs;
                    nearfield::halfturn_parts(fmaxf(r2, 0x1p-120f), c, s, rs);
                    rs = r2 > fl2 ? rs : 0.0f;
                    const float wr = c * rs, wi = s * rs;
                    ar[r] = fmaf(wr, qj.x, ar[r]);
                    ar[r] = fmaf(-wi, qj.y, ar[r]);
                    ai[r] = fmaf(wr, qj.y, ai[r]);
                    ai[r] = fmaf(wi, qj.x, ai[r]);
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                sr[r] += (double)ar[r];
                si[r] += (double)ai[r];
            }
            __syncthreads();
        }
        for (int off = L; off < 64; off <<= 1)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                sr[r] += __shfl_xor(sr[r], off);
                si[r] += __shfl_xor(si[r], off);
            }
        if (g == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int ti = li + r * L;
                if (ti < tc) {
                    const int it = tstart + t0 + ti;
                    double2 f = a.far[it];
                    f.x = fma(sr[r], a.wscale, f.x);
                    f.y = fma(si[r], a.wscale, f.y);
                    if (a.out_mode & 1) f.y = -f.y;   // k < 0: conjugate of the |k| result
                    if (a.out_mode & 2) f.y = 0.0;    // k == 0 with real charges: the result is real
                    a.out[a.perm[it]] = f;
                }
            }
        }
    }
}

template <bool PH, bool LOWK, int R>
__device__ __forceinline__ void p2p_mixed_body(const P2PArgs& a, const int leaf, P2PStage& st) {
    if constexpr (!PH && !LOWK) {
        if (a.coord32 && a.coord32[leaf]) {   // uniform: one leaf per workgroup
            p2p_mixed_body_c32<R>(a, leaf, st);
            return;
        }
    }
    double2* sxy = st.mixed.xy;
    double2* szq = st.mixed.zq;
    const int b = a.leaf_box[leaf];
    const int tstart = a.bx.pstart[b], tcnt = a.bx.count[b];
    const double sc = a.scale;
    const double cx = a.bx.centre[3 * b], cy = a.bx.centre[3 * b + 1], cz = a.bx.centre[3 * b + 2];
    const int r0 = a.rng_ptr[leaf], r1 = a.rng_ptr[leaf + 1];
    const int S = a.leaf_src_total[leaf];
    // the floor in scaled coordinates, and at least 2^-120 so that an FP32 r^2 above it is
    // a normal number
    const double fl2 = fmax(a.floor2[leaf] * sc * sc, 0x1p-120);
    const float kf = (float)a.k;
    const int lane = threadIdx.x;
    for (int t0 = 0, tc = 0; t0 < tcnt; t0 += tc) {
        tc = p2p_chunk<R>(tcnt - t0);
        const int need = (tc + R - 1) / R;
        int L = 1;
        while (L < need) L <<= 1;
        const int G = 64 / L;
        const int li = lane & (L - 1), g = lane / L;
        double xt[R], yt[R], zt[R], sr[R], si[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int it = tstart + t0 + min(li + r * L, tc - 1);
            xt[r] = (a.X[it] - cx) * sc;
            yt[r] = (a.Y[it] - cy) * sc;
            zt[r] = (a.Z[it] - cz) * sc;
            sr[r] = 0.0;
            si[r] = 0.0;
        }
        int ri = r0;
        for (int base = 0; base < S; base += 64) {
            const int v = base + lane;
            if (v < S) {
                const int idx = p2p_source_index(v, r1, a.rng_start, a.rng_pref, ri);
                const double2 qq = a.q[idx];
                const float2 qf = make_float2((float)qq.x, (float)qq.y);
                sxy[lane] = make_double2((a.X[idx] - cx) * sc, (a.Y[idx] - cy) * sc);
                szq[lane] = make_double2((a.Z[idx] - cz) * sc, __builtin_bit_cast(double, qf));
            }
            __syncthreads();
            const int nb = min(64, S - base);
            float ar[R], ai[R];
#pragma unroll
            for (int r = 0; r < R; ++r) ar[r] = ai[r] = 0.0f;
            for (int j = g; j < nb; j += G) {
                const double2 pxy = sxy[j], pzq = szq[j];
                const float2 qj = __builtin_bit_cast(float2, pzq.y);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const double dx = xt[r] - pxy.x, dy = yt[r] - pxy.y, dz = zt[r] - pzq.x;
                    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                    float wr, wi;
                    if constexpr (LOWK) {
                        const bool ok = r2 > fl2;
                        const float rs = nearfield::rsq((float)r2);
                        wr = ok ? rs : 0.0f;
                        wi = ok ? kf : 0.0f;
                    } else {
                        nearfield::pair_weight<PH>(r2, fl2, wr, wi);
                    }
                    ar[r] = fmaf(wr, qj.x, ar[r]);
                    ar[r] = fmaf(-wi, qj.y, ar[r]);
                    ai[r] = fmaf(wr, qj.y, ai[r]);
                    ai[r] = fmaf(wi, qj.x, ai[r]);
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                sr[r] += (double)ar[r];
                si[r] += (double)ai[r];
            }
            __syncthreads();
        }
        for (int off = L; off < 64; off <<= 1)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                sr[r] += __shfl_xor(sr[r], off);
                si[r] += __shfl_xor(si[r], off);
            }
        if (g == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int ti = li + r * L;
                if (ti < tc) {
                    const int it = tstart + t0 + ti;
                    double2 f = a.far[it];
                    f.x = fma(sr[r], a.wscale, f.x);
                    f.y = fma(si[r], a.wscale, f.y);
                    if (a.out_mode & 1) f.y = -f.y;   // k < 0: conjugate of the |k| result
                    if (a.out_mode & 2) f.y = 0.0;    // k == 0 with real charges: the result is real
                    a.out[a.perm[it]] = f;
                }
            }
        }
    }
}

// Leaf of workgroup i of a launch over n leaves: workgroup i runs on XCD i % 8, and XCD x
// takes the contiguous range of n / 8 (+1 for x < n % 8) leaves starting at
// x (n / 8) + min(x, n % 8), so that its L2 holds one region's points (leaves are in Morton
// order, and a leaf's sources are its neighbours).  A bijection of [0, n).
__device__ inline int xcd_leaf(int i, int n) {
    const int q = n / kXcds, r = n % kXcds, x = i % kXcds;
    return x * q + min(x, r) + i / kXcds;
}

template <bool PH, bool LOWK, int R>
__global__ __launch_bounds__(64) void p2p_mixed_kernel(P2PArgs a) {
    __shared__ P2PStage st;
    p2p_mixed_body<PH, LOWK, R>(a, a.leaf0 + xcd_leaf(blockIdx.x, gridDim.x), st);
}

// FP64 near field (the plain form): one wave per target leaf; targets of a chunk of up to
// 64 points are spread over lane groups: with T' = next power of two >= chunk size, lane
// l holds target l % T' and scans the sources j == l / T' (mod 64/T') of each 64-point
// source batch staged in LDS (coordinates relative to the target leaf centre).  Lane-group
// partial sums are combined with wavefront shuffles.  Sources of a leaf are a list of
// point ranges (adjacent leaves, and the W/X interactions evaluated directly); pairs at or
// below the leaf's coincidence floor are skipped.
__device__ __forceinline__ void p2p_fp64_body(const P2PArgs& a, const int leaf, P2PStage& st) {
    double *sx = st.fp64.x, *sy = st.fp64.y, *sz = st.fp64.z, *sqr = st.fp64.qr, *sqi = st.fp64.qi;
    const int b = a.leaf_box[leaf];
    const int tstart = a.bx.pstart[b], tcnt = a.bx.count[b];
    const double cx = a.bx.centre[3 * b], cy = a.bx.centre[3 * b + 1], cz = a.bx.centre[3 * b + 2];
    const int r0 = a.rng_ptr[leaf], r1 = a.rng_ptr[leaf + 1];
    const int S = a.leaf_src_total[leaf];
    const double fl2 = a.floor2[leaf];
    const int lane = threadIdx.x;
    // target chunks of 64; a remainder of 33..63 targets goes as 32 + the rest, so that the
    // lanes are not left idle with T' = 64 (each lane then scans half or less of the sources)
    for (int t0 = 0, tc = 0; t0 < tcnt; t0 += tc) {
        tc = tcnt - t0 >= 64 ? 64 : min(32, tcnt - t0);
        int Tp = 1;
        while (Tp < tc) Tp <<= 1;
        const int G = 64 / Tp;
        const int ti = lane & (Tp - 1), g = lane / Tp;
        const bool has_t = ti < tc;
        const int it = tstart + t0 + (has_t ? ti : 0);
        const double xt = a.X[it] - cx, yt = a.Y[it] - cy, zt = a.Z[it] - cz;
        double sr = 0.0, si = 0.0;
        int ri = r0;
        for (int base = 0; base < S; base += 64) {
            const int v = base + lane;
            if (v < S) {
                const int idx = p2p_source_index(v, r1, a.rng_start, a.rng_pref, ri);
                sx[lane] = a.X[idx] - cx;
                sy[lane] = a.Y[idx] - cy;
                sz[lane] = a.Z[idx] - cz;
                const double2 qq = a.q[idx];
                sqr[lane] = qq.x;
                sqi[lane] = qq.y;
            }
            __syncthreads();
            const int nb = min(64, S - base);
            for (int j = g; j < nb; j += G) {
                const double dx = xt - sx[j], dy = yt - sy[j], dz = zt - sz[j];
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > fl2) {
                    double wr, wi;
                    P2PMath<double>::kernel(r2, a.k, wr, wi);
                    const double qr = sqr[j], qi = sqi[j];
                    sr += wr * qr - wi * qi;
                    si += wr * qi + wi * qr;
                }
            }
            __syncthreads();
        }
        for (int off = Tp; off < 64; off <<= 1) {
            sr += __shfl_xor(sr, off);
            si += __shfl_xor(si, off);
        }
        if (g == 0 && has_t) {
            double2 f = a.far[it];
            f.x += sr;
            f.y += si;
            if (a.out_mode & 1) f.y = -f.y;   // k < 0: conjugate of the |k| result
            if (a.out_mode & 2) f.y = 0.0;    // k == 0 with real charges: the result is real
            a.out[a.perm[it]] = f;
        }
    }
}

__global__ __launch_bounds__(64) void p2p_fp64_kernel(P2PArgs a) {
    __shared__ P2PStage st;
    p2p_fp64_body(a, a.leaf0 + xcd_leaf(blockIdx.x, gridDim.x), st);
}

// Persistent near field over the leaves of many problems (batches), one launch per leaf
// class CLS (kP2PClass*): one wave per workgroup takes (problem, leaf) work from the
// queues q until all are empty; args[w.x] is the problem's near-field description.
template <int CLS>
__global__ __launch_bounds__(64) void p2p_persistent_kernel(const P2PBatchArgs* __restrict__ args, WorkQueues q) {
    __shared__ P2PStage st;
    const int home = blockIdx.x % kXcds;
    for (;;) {
        int w = 0;
        if (threadIdx.x == 0) w = work_dequeue(q, home);
        w = __builtin_amdgcn_readfirstlane(w);   // uniform: the descriptor is read with scalar loads
        if (w < 0) break;
        const int2 wk = q.work[w];
        const int prob = __builtin_amdgcn_readfirstlane(wk.x), leaf = __builtin_amdgcn_readfirstlane(wk.y);
        const P2PArgs a = args[prob].a;
        if constexpr (CLS == kP2PClassFp64)
            p2p_fp64_body(a, leaf, st);
        else if constexpr (CLS == kP2PClassLowK)
            p2p_mixed_body<false, true, kP2PBatchTargetsPerLane>(a, leaf, st);
        else if constexpr (CLS == kP2PClassPhase32)
            p2p_mixed_body<false, false, kP2PBatchTargetsPerLane>(a, leaf, st);
        else
            p2p_mixed_body<true, false, kP2PBatchTargetsPerLane>(a, leaf, st);
        __syncthreads();
    }
}

// P2P with the field: the potential and E = -grad phi of the near field (same
// decomposition and source ranges as p2p_kernel), added to the far-field values.
// field[3 i + a] holds component a of point i (input order).
template <typename T, bool PH>
__global__ __launch_bounds__(64) void p2p_field_kernel(int leaf0, const int* leaf_box, DevBoxes bx,
                                                       const int* rng_ptr, const int* rng_start, const int* rng_pref,
                                                       const int* leaf_src_total, const double* floor2,
                                                       const double* X, const double* Y, const double* Z,
                                                       const double2* q, double k, const double2* far,
                                                       const double2* far_e, const int* perm, int out_mode,
                                                       double2* out, double2* field) {
    // coordinates relative to the leaf centre and r^2 in FP64 at every precision: FP32
    // coordinates would carry an absolute error of 2^-24 times the leaf size, a large
    // relative error of r for close pairs (points on surfaces); FP64 adds and FMAs issue
    // at the rate of the (unpacked) FP32 ones
    using C = double;
    __shared__ C sx[64], sy[64], sz[64];
    __shared__ T sqr[64], sqi[64];
    const int leaf = leaf0 + blockIdx.x;
    const int b = leaf_box[leaf];
    const int tstart = bx.pstart[b], tcnt = bx.count[b];
    const double cx = bx.centre[3 * b], cy = bx.centre[3 * b + 1], cz = bx.centre[3 * b + 2];
    const int r0 = rng_ptr[leaf], r1 = rng_ptr[leaf + 1];
    const int S = leaf_src_total[leaf];
    const double fl2 = floor2[leaf];
    const int lane = threadIdx.x;
    const T kk = (T)k;
    // target chunks of 64; a remainder of 33..63 targets goes as 32 + the rest, so that the
    // lanes are not left idle with T' = 64 (each lane then scans half or less of the sources)
    for (int t0 = 0, tc = 0; t0 < tcnt; t0 += tc) {
        tc = tcnt - t0 >= 64 ? 64 : min(32, tcnt - t0);
        int Tp = 1;
        while (Tp < tc) Tp <<= 1;
        const int G = 64 / Tp;
        const int ti = lane & (Tp - 1), g = lane / Tp;
        const bool has_t = ti < tc;
        const int it = tstart + t0 + (has_t ? ti : 0);
        const C xt = (C)(X[it] - cx), yt = (C)(Y[it] - cy), zt = (C)(Z[it] - cz);
        double sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // phi, Ex, Ey, Ez (re, im); FP64 over batches
        for (int base = 0; base < S; base += 64) {
            T acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            const int v = base + lane;
            if (v < S) {
                int lo = r0, hi = r1 - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (rng_pref[mid] <= v)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                const int idx = rng_start[lo] + (v - rng_pref[lo]);
                sx[lane] = (C)(X[idx] - cx);
                sy[lane] = (C)(Y[idx] - cy);
                sz[lane] = (C)(Z[idx] - cz);
                const double2 qq = q[idx];
                sqr[lane] = (T)qq.x;
                sqi[lane] = (T)qq.y;
            }
            __syncthreads();
            const int nb = min(64, S - base);
            for (int j = g; j < nb; j += G) {
                const C dx = xt - sx[j], dy = yt - sy[j], dz = zt - sz[j];
                const C r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > fl2) {
                    T wr, wi, fr, fi;
                    if constexpr (PH) {
                        float rinv, kr;
                        radian_weight_phase64(r2, k, wr, wi, rinv, kr);
                        const float rinv2 = rinv * rinv;
                        fr = (wr + kr * wi) * rinv2;
                        fi = (wi - kr * wr) * rinv2;
                    } else {
                        field_weights<T>((T)r2, kk, wr, wi, fr, fi);
                    }
                    const T qr = sqr[j], qi = sqi[j];
                    acc[0] += wr * qr - wi * qi;
                    acc[1] += wr * qi + wi * qr;
                    const T gr = fr * qr - fi * qi, gi = fr * qi + fi * qr;
                    acc[2] += (T)dx * gr;
                    acc[3] += (T)dx * gi;
                    acc[4] += (T)dy * gr;
                    acc[5] += (T)dy * gi;
                    acc[6] += (T)dz * gr;
                    acc[7] += (T)dz * gi;
                }
            }
#pragma unroll
            for (int c = 0; c < 8; ++c) sum[c] += (double)acc[c];
            __syncthreads();
        }
        for (int off = Tp; off < 64; off <<= 1)
#pragma unroll
            for (int c = 0; c < 8; ++c) sum[c] += __shfl_xor(sum[c], off);
        if (g == 0 && has_t) {
            const int o = perm[it];
            double2 f = far[it];
            f.x += sum[0];
            f.y += sum[1];
            if (out_mode & 1) f.y = -f.y;   // k < 0: conjugate of the |k| result
            if (out_mode & 2) f.y = 0.0;    // k == 0 with real charges: the result is real
            out[o] = f;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                double2 e = far_e[3 * (int64_t)it + a];
                e.x += sum[2 + 2 * a];
                e.y += sum[3 + 2 * a];
                if (out_mode & 1) e.y = -e.y;
                if (out_mode & 2) e.y = 0.0;
                field[3 * (int64_t)o + a] = e;
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// FP64 direct sum at sampled targets: grid (target tiles of 64, source chunks);
// partial sums combined with FP64 atomics into a zeroed output.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void direct_kernel(const double* X, const double* Y, const double* Z,
                                                     const double2* q, int64_t n, double k, const double* tx,
                                                     const double* ty, const double* tz, int64_t nt,
                                                     int64_t chunk, double2* out) {
    __shared__ double sx[256], sy[256], sz[256], sqr[256], sqi[256];
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool valid = t < nt;
    const double x = valid ? tx[t] : 0.0, y = valid ? ty[t] : 0.0, z = valid ? tz[t] : 0.0;
    const int64_t s0 = (int64_t)blockIdx.y * chunk, s1 = min(n, s0 + chunk);
    double ar = 0.0, ai = 0.0;
    for (int64_t base = s0; base < s1; base += 256) {
        const int64_t v = base + threadIdx.x;
        if (v < s1) {
            sx[threadIdx.x] = X[v];
            sy[threadIdx.x] = Y[v];
            sz[threadIdx.x] = Z[v];
            sqr[threadIdx.x] = q[v].x;
            sqi[threadIdx.x] = q[v].y;
        }
        __syncthreads();
        const int nb = (int)min((int64_t)256, s1 - base);
        for (int j = 0; j < nb; ++j) {
            const double dx = x - sx[j], dy = y - sy[j], dz = z - sz[j];
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (r > 0.0) {
                double s, c;
                sincos(k * r, &s, &c);
                const double wr = c / r, wi = s / r;
                ar += wr * sqr[j] - wi * sqi[j];
                ai += wr * sqi[j] + wi * sqr[j];
            }
        }
        __syncthreads();
    }
    if (valid) {
        atomicAdd(&out[t].x, ar);
        atomicAdd(&out[t].y, ai);
    }
}

}  // namespace

// ---------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------
void upload_legendre_coefficients() {
    // constant memory is per device: upload once on every device used
    constexpr int kMaxDevices = 64;
    static bool done[kMaxDevices] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) fail(-3, "hipGetDevice failed");
    if (done[dev]) return;
    LegCoef h{};
    for (int n = 0; n < kMillerN; ++n) h.miller[n] = 1.0 / ((2.0 * n + 1.0) * (2.0 * n + 3.0));
    for (int m = 0; m <= kMaxP; ++m) {
        h.de[m] = make_double2(m > 0 ? std::sqrt((2.0 * m + 1.0) / (2.0 * m)) : 0.0, std::sqrt(2.0 * m + 3.0));
        h.cm[m] = m > 0 ? -h.de[m].x * h.cm[m - 1] : 0.28209479177387814;   // 1/sqrt(4 pi)
        for (int n = 0; n <= kMaxP; ++n) {
            if (n >= m + 2)
                h.ab[m][n] = make_double2(
                    std::sqrt((4.0 * n * n - 1.0) / ((double)n * n - (double)m * m)),
                    std::sqrt(((n - 1.0) * (n - 1.0) - (double)m * m) / (4.0 * (n - 1.0) * (n - 1.0) - 1.0)));
            else
                h.ab[m][n] = make_double2(0.0, 0.0);
        }
    }
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_leg), &h, sizeof(h)) != hipSuccess)
        fail(-3, "hipMemcpyToSymbol(Legendre coefficients) failed");
    done[dev] = true;
}

static void check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(-3, std::string("kernel launch failed (") + what + "): " + hipGetErrorString(e));
}

void launch_fill_zero(double2* buf, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, buf, n);
    check_launch("fill_zero");
}

void launch_gather_charges(const void* q, int q_type, const int* perm, int64_t n, int conj, double2* q_sorted,
                           hipStream_t s) {
    hipLaunchKernelGGL(gather_charges_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, q, q_type, perm,
                       n, conj, q_sorted);
    check_launch("gather_charges");
}

int p2m_chunk(int pmax) {
    // LDS with the coefficients within 22 KB for 128 threads (7 workgroups per CU) and 36 KB
    // for 256 (4 per CU)
    const long budget =
        (p2m_threads(pmax) == 128 ? 22 : 36) * 1024L - (long)leg_lds_doubles(pmax) * (long)sizeof(double);
    return (int)std::max<long>(1, std::min<long>(64, budget / (long)p2m_bytes_per_point(pmax)));
}

size_t p2m_lds_bytes(int pmax, int chunk) {
    // the chunk region also holds the point groups' partial sums at the end of a leaf:
    // (G - 1) x blocks x 4 x 32 bytes, with G x blocks <= threads
    return (size_t)leg_lds_doubles(pmax) * sizeof(double) +
           std::max<size_t>((size_t)chunk * p2m_bytes_per_point(pmax),
                            (size_t)p2m_threads(pmax) * kP2MBlockN * 2 * sizeof(double2));
}

// resident workgroups of a persistent leaf loop: the occupancy of (fn, threads, shm) on
// every CU, a multiple of kXcds, at most the leaves rounded up to kXcds
static int leaf_loop_grid(const void* fn, int threads, size_t shm, int n_items) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, shm))
        fail(-3, "occupancy query (leaf loop) failed");
    const int full = std::max(kXcds, std::max(1, per_cu) * cus / kXcds * kXcds);
    return std::min(full, (n_items + kXcds - 1) / kXcds * kXcds);
}

void launch_p2m(const LeafDesc* leaves, int n_items, int chunk, DevLevels lv, const double* X, const double* Y,
                const double* Z, const double2* q, double k, int pmax, double2* M, hipStream_t s) {
    if (n_items <= 0) return;
    upload_legendre_coefficients();
    const size_t shm = p2m_lds_bytes(pmax, chunk);
    if (p2m_threads(pmax) == 128) {
        const int grid = leaf_loop_grid((const void*)p2m_kernel<128>, 128, shm, n_items);
        hipLaunchKernelGGL(p2m_kernel<128>, dim3(grid), dim3(128), shm, s, leaves, n_items, chunk, pmax, lv, X, Y,
                           Z, q, k, M);
    } else {
        const int grid = leaf_loop_grid((const void*)p2m_kernel<256>, 256, shm, n_items);
        hipLaunchKernelGGL(p2m_kernel<256>, dim3(grid), dim3(256), shm, s, leaves, n_items, chunk, pmax, lv, X, Y,
                           Z, q, k, M);
    }
    check_launch("p2m");
}

void launch_l2p(const LeafDesc* leaves, int n_items, DevLevels lv, const double* X, const double* Y, const double* Z,
                const double2* L, double k, int pmax, double2* far, hipStream_t s) {
    if (n_items <= 0) return;
    upload_legendre_coefficients();
    const size_t shm = (size_t)l2p_lds_doubles(pmax) * sizeof(double);
    const int grid = leaf_loop_grid((const void*)l2p_kernel, 64, shm, n_items);
    hipLaunchKernelGGL(l2p_kernel, dim3(grid), dim3(64), shm, s, leaves, n_items, pmax, lv, X, Y, Z, L, k, far);
    check_launch("l2p");
}

size_t l2p_field_lds_bytes(int p) {
    const size_t P1 = p + 1, P2 = p + 2;
    return (P1 * P1 + 3 * P2 * P2) * sizeof(double2) + 64 * P2 * sizeof(double);
}

void launch_l2p_field(const int* leaves, int n_items, DevBoxes bx, DevLevels lv, const double* X, const double* Y,
                      const double* Z, const double2* L, double k, int pmax, double2* far, double2* far_e,
                      hipStream_t s) {
    if (n_items <= 0) return;
    upload_legendre_coefficients();
    if (pmax + 1 > kMaxP) fail(-2, "the field needs leaf degrees of at most 39");
    const size_t shm = l2p_field_lds_bytes(pmax);
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void*)l2p_field_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)l2p_field_lds_bytes(kMaxP - 1));
        if (e != hipSuccess) fail(-3, std::string("hipFuncSetAttribute(l2p_field) failed: ") + hipGetErrorString(e));
        attr = true;
    }
    hipLaunchKernelGGL(l2p_field_kernel, dim3(n_items), dim3(64), shm, s, leaves, bx, lv, X, Y, Z, L, k, far, far_e);
    check_launch("l2p_field");
}

void launch_p2p_field(int precision, int n_leaves, int n_fp32, const int* leaf_box, DevBoxes bx,
                      const int* rng_ptr, const int* rng_start, const int* rng_pref, const int* leaf_src_total,
                      const double* floor2, const double* X, const double* Y, const double* Z, const double2* q,
                      double k, const double2* far, const double2* far_e, const int* perm, int out_mode,
                      double2* out, double2* field, hipStream_t s) {
    if (n_leaves <= 0) return;
    if (precision == 0) {
        hipLaunchKernelGGL((p2p_field_kernel<double, false>), dim3(n_leaves), dim3(64), 0, s, 0, leaf_box, bx,
                           rng_ptr, rng_start, rng_pref, leaf_src_total, floor2, X, Y, Z, q, k, far, far_e, perm,
                           out_mode, out, field);
    } else {
        if (n_fp32 > 0)
            hipLaunchKernelGGL((p2p_field_kernel<float, false>), dim3(n_fp32), dim3(64), 0, s, 0, leaf_box, bx,
                               rng_ptr, rng_start, rng_pref, leaf_src_total, floor2, X, Y, Z, q, k, far, far_e,
                               perm, out_mode, out, field);
        if (n_leaves > n_fp32)
            hipLaunchKernelGGL((p2p_field_kernel<float, true>), dim3(n_leaves - n_fp32), dim3(64), 0, s, n_fp32,
                               leaf_box, bx, rng_ptr, rng_start, rng_pref, leaf_src_total, floor2, X, Y, Z, q, k,
                               far, far_e, perm, out_mode, out, field);
    }
    check_launch("p2p_field");
}

void launch_xlate(const XlateArgs& a, int n_targets, hipStream_t s) {
    if (n_targets <= 0) return;
    if (a.split < 1 || (a.split > 1 && !a.accumulate)) fail(-9, "xlate: split needs accumulate");
    const int nmx = std::max((a.ps + 1) * (a.ps + 1), (a.pt + 1) * (a.pt + 1));
    const size_t shm = 2 * (size_t)nmx * sizeof(double2) + (kMaxP + 1) * sizeof(int);
    hipLaunchKernelGGL(xlate_kernel, dim3((unsigned)n_targets * (unsigned)a.split), dim3(kXlateThreads), shm, s, a);
    check_launch("xlate");
}

P2PArgs p2p_scaled(P2PArgs a, int precision, bool lowk) {
    if (precision != 0) {
        // coordinates in units of pi / k: the weights are e^{i pi u}/u with u = k r / pi
        a.scale = lowk ? 1.0 : a.k / M_PI;
        a.wscale = lowk ? 1.0 : a.k / M_PI;
    }
    return a;
}

template <int CLS>
static const void* p2p_persistent_fn() {
    return (const void*)p2p_persistent_kernel<CLS>;
}

static const void* p2p_persistent_class_fn(int cls) {
    switch (cls) {
        case kP2PClassFp64: return p2p_persistent_fn<kP2PClassFp64>();
        case kP2PClassLowK: return p2p_persistent_fn<kP2PClassLowK>();
        case kP2PClassPhase32: return p2p_persistent_fn<kP2PClassPhase32>();
        default: return p2p_persistent_fn<kP2PClassPhase64>();
    }
}

int p2p_persistent_grid(int cls) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, p2p_persistent_class_fn(cls), 64, 0))
        fail(-3, "occupancy query (p2p persistent) failed");
    return std::max(kXcds, (std::max(1, per_cu) * cus) / kXcds * kXcds);
}

void launch_p2p_persistent(int cls, const P2PBatchArgs* d_args, const WorkQueues& q, int grid, hipStream_t s) {
    if (grid <= 0) return;
    switch (cls) {
        case kP2PClassFp64:
            hipLaunchKernelGGL(p2p_persistent_kernel<kP2PClassFp64>, dim3(grid), dim3(64), 0, s, d_args, q);
            break;
        case kP2PClassLowK:
            hipLaunchKernelGGL(p2p_persistent_kernel<kP2PClassLowK>, dim3(grid), dim3(64), 0, s, d_args, q);
            break;
        case kP2PClassPhase32:
            hipLaunchKernelGGL(p2p_persistent_kernel<kP2PClassPhase32>, dim3(grid), dim3(64), 0, s, d_args, q);
            break;
        default:
            hipLaunchKernelGGL(p2p_persistent_kernel<kP2PClassPhase64>, dim3(grid), dim3(64), 0, s, d_args, q);
    }
    check_launch("p2p persistent");
}

__global__ void clear_ints_kernel(int* p, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

void launch_clear_ints(int* p, int n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(clear_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    check_launch("clear");
}

template <int R>
static void launch_p2p_mixed(P2PArgs a, int n_leaves, int n_fp32, bool lowk, hipStream_t s) {
    a = p2p_scaled(a, 1, lowk);
    if (lowk) {
        hipLaunchKernelGGL((p2p_mixed_kernel<false, true, R>), dim3(n_leaves), dim3(64), 0, s, a);
        return;
    }
    if (n_fp32 > 0) hipLaunchKernelGGL((p2p_mixed_kernel<false, false, R>), dim3(n_fp32), dim3(64), 0, s, a);
    if (n_leaves > n_fp32) {
        a.leaf0 = n_fp32;
        hipLaunchKernelGGL((p2p_mixed_kernel<true, false, R>), dim3(n_leaves - n_fp32), dim3(64), 0, s, a);
    }
}

void launch_p2p(const P2PArgs& a0, int precision, int n_leaves, int n_fp32, bool lowk, int tile, hipStream_t s) {
    if (n_leaves <= 0) return;
    P2PArgs a = a0;
    a.leaf0 = 0;
    if (precision == 0)
        hipLaunchKernelGGL(p2p_fp64_kernel, dim3(n_leaves), dim3(64), 0, s, a);
    else if (tile == 1)
        launch_p2p_mixed<1>(a, n_leaves, n_fp32, lowk, s);
    else if (tile == 4)
        launch_p2p_mixed<4>(a, n_leaves, n_fp32, lowk, s);
    else
        launch_p2p_mixed<2>(a, n_leaves, n_fp32, lowk, s);
    check_launch("p2p");
}

// One wave per leaf: lane t holds target t of a 64-target chunk; every source of the leaf is
// read by all lanes (uniform loads), FP64 throughout.
__global__ __launch_bounds__(64) void p2p_coord_check_kernel(P2PArgs a, double tol, unsigned char* flag) {
    const int leaf = blockIdx.x;
    const int b = a.leaf_box[leaf];
    const int tstart = a.bx.pstart[b], tcnt = a.bx.count[b];
    const double cx = a.bx.centre[3 * b], cy = a.bx.centre[3 * b + 1], cz = a.bx.centre[3 * b + 2];
    const int r0 = a.rng_ptr[leaf], r1 = a.rng_ptr[leaf + 1];
    const int S = a.leaf_src_total[leaf];
    const double fl2 = a.floor2[leaf];
    const int lane = threadIdx.x;
    double cmax = 0.0, ratio = 0.0;
    for (int t0 = 0; t0 < tcnt; t0 += 64) {
        const bool has = t0 + lane < tcnt;
        const int it = tstart + t0 + (has ? lane : 0);
        const double xt = a.X[it] - cx, yt = a.Y[it] - cy, zt = a.Z[it] - cz;
        if (has) cmax = fmax(cmax, fmax(fabs(xt), fmax(fabs(yt), fabs(zt))));
        double s2 = 0.0, s4 = 0.0;
        int ri = r0;
        for (int v = 0; v < S; ++v) {
            const int idx = p2p_source_index(v, r1, a.rng_start, a.rng_pref, ri);
            const double xs = a.X[idx] - cx, ys = a.Y[idx] - cy, zs = a.Z[idx] - cz;
            if (t0 == 0) cmax = fmax(cmax, fmax(fabs(xs), fmax(fabs(ys), fabs(zs))));
            const double dx = xt - xs, dy = yt - ys, dz = zt - zs;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > fl2) {
                const double w = 1.0 / r2;
                s2 += w;
                s4 += w * w;
            }
        }
        if (has && s2 > 0.0) ratio = fmax(ratio, sqrt(s4 / s2));
    }
    for (int off = 1; off < 64; off <<= 1) {
        cmax = fmax(cmax, __shfl_xor(cmax, off));
        ratio = fmax(ratio, __shfl_xor(ratio, off));
    }
    if (lane == 0) flag[leaf] = 0x1p-21 * cmax * ratio <= tol ? 1 : 0;
}

void launch_p2p_coord_check(const P2PArgs& a, int n, double tol, unsigned char* flag, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(p2p_coord_check_kernel, dim3(n), dim3(64), 0, s, a, tol, flag);
    check_launch("p2p coordinate check");
}

void launch_direct(const double* X, const double* Y, const double* Z, const double2* q, int64_t n, double k,
                   const double* tx, const double* ty, const double* tz, int64_t nt, double2* out, hipStream_t s) {
    if (nt <= 0) return;
    const int64_t tiles = (nt + 255) / 256;
    int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>(4096 / tiles, (n + 4095) / 4096));
    const int64_t chunk = (n + nchunk - 1) / nchunk;
    nchunk = (n + chunk - 1) / chunk;
    launch_fill_zero(out, nt, s);
    hipLaunchKernelGGL(direct_kernel, dim3((unsigned)tiles, (unsigned)nchunk), dim3(256), 0, s, X, Y, Z, q, n, k, tx,
                       ty, tz, nt, chunk, out);
    check_launch("direct");
}

}  // namespace hfmm
