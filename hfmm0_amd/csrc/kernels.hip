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
// Also the downward-recurrence factors of jhat, miller[n] = 1 / ((2n+1)(2n+3)).
constexpr int kMillerN = 512;
struct LegCoef {
    double a[kMaxP + 1][kMaxP + 1];
    double b[kMaxP + 1][kMaxP + 1];
    double d[kMaxP + 1];
    double e[kMaxP + 1];
    double miller[kMillerN];
};
__constant__ LegCoef c_leg;

// jhat_n(x) for n = 0..p by Miller's downward recurrence
//   jhat_{n-1} = jhat_n - x^2 jhat_{n+1} / ((2n+1)(2n+3)),
// normalised with jhat_0 = sin(x)/x (or jhat_1 = 3(sin x - x cos x)/x^3 when that is
// the larger one).  out[n * stride].  The recurrence starts at N = p + 1.5 xmax + 12 for
// an upper bound xmax >= x common to the caller's workgroup (k R of the box), so every lane
// walks the same n and the factor table is read at a uniform index.  Against SciPy's
// spherical Bessel functions (p = 2..40, x = 1e-3..150) a start of p + 1.5 x + 8 already
// reaches the rounding floor (7e-14 relative to the neighbouring orders' magnitude); + 6
// gives 2e-13, + 4 gives 9e-11.
__device__ void jhat_dev(double x, int p, double* out, int stride, double xmax) {
    x = fabs(x);
    if (x == 0.0) {
        for (int n = 0; n <= p; ++n) out[n * stride] = 1.0;
        return;
    }
    const int N = p + (int)(1.5 * fmax(xmax, x)) + 12;
    double a_next = 0.0, a = 1.0;
    const double x2 = x * x;
    for (int n = N; n >= 1; --n) {
        const double f = n < kMillerN ? c_leg.miller[n] : 1.0 / ((2.0 * n + 1.0) * (2.0 * n + 3.0));
        const double am1 = a - x2 * a_next * f;
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
// P2M: one workgroup per leaf; points processed in chunks staged in LDS.
//  1. one thread per point: offsets, angles and the radial factors
//     q_j jhat_n(k rho_j) (rho_j / R)^n;
//  2. one thread per (point j, order m): the normalised Legendre column
//     P_n^m(cos beta_j), n = m..p, and e^{-i m alpha_j} (powers of e^{-i alpha_j});
//  3. one thread per coefficient (n, m): accumulation over the chunk's points.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void p2m_kernel(const int* leaves, int chunk, DevBoxes bx, DevLevels lv,
                                                  const double* X, const double* Y, const double* Z,
                                                  const double2* q, double k, double2* M) {
    extern __shared__ double smem[];
    const int b = leaves[blockIdx.x];
    const int L = bx.level[b];
    const int p = lv.p[L];
    const double R = lv.R[L];
    const int P1 = p + 1, tri = P1 * (P1 + 1) / 2, nc = P1 * P1;
    double2* s_qrad = (double2*)smem;             // chunk x P1
    double2* s_ph = s_qrad + chunk * P1;          // chunk x P1
    double2* s_geo = s_ph + chunk * P1;           // chunk x 2: (cos b, sin b), (cos a, sin a)
    double* s_leg = (double*)(s_geo + 2 * chunk); // chunk x tri
    const double cx = bx.centre[3 * b], cy = bx.centre[3 * b + 1], cz = bx.centre[3 * b + 2];
    const int ps = bx.pstart[b], cnt = bx.count[b];
    double2 acc[kAccPerThread];
#pragma unroll
    for (int t = 0; t < kAccPerThread; ++t) acc[t] = make_double2(0.0, 0.0);

    for (int j0 = 0; j0 < cnt; j0 += chunk) {
        const int nch = min(chunk, cnt - j0);
        for (int j = threadIdx.x; j < nch; j += blockDim.x) {
            const int i = ps + j0 + j;
            const double dx = X[i] - cx, dy = Y[i] - cy, dz = Z[i] - cz;
            const double rxy2 = dx * dx + dy * dy;
            const double rho = sqrt(rxy2 + dz * dz), rxy = sqrt(rxy2);
            s_geo[2 * j] = rho > 0.0 ? make_double2(dz / rho, rxy / rho) : make_double2(1.0, 0.0);
            s_geo[2 * j + 1] = rxy > 0.0 ? make_double2(dx / rxy, dy / rxy) : make_double2(1.0, 0.0);
            double* rad = &s_qrad[j * P1].x;
            jhat_dev(k * rho, p, rad, 2, k * R);
            const double2 qq = q[i];
            const double rr = rho / R;
            double pw = 1.0;
            for (int n = 0; n <= p; ++n) {
                const double g = rad[2 * n] * pw;
                s_qrad[j * P1 + n] = make_double2(qq.x * g, qq.y * g);
                pw *= rr;
            }
        }
        __syncthreads();
        for (int w = threadIdx.x; w < nch * P1; w += blockDim.x) {
            const int j = w / P1, m = w - j * P1;
            const double2 cb = s_geo[2 * j], ea = s_geo[2 * j + 1];
            const double ct = cb.x, st = cb.y;
            double pmm = 0.28209479177387814;   // 1/sqrt(4 pi)
            double2 z = make_double2(1.0, 0.0);
            for (int i2 = 1; i2 <= m; ++i2) {
                pmm *= -c_leg.d[i2] * st;
                z = cmul(z, ea);
            }
            s_ph[j * P1 + m] = make_double2(z.x, -z.y);
            double* Pl = s_leg + j * tri;
            Pl[m * (m + 1) / 2 + m] = pmm;
            if (m < p) {
                double p2 = pmm, p1 = c_leg.e[m] * ct * pmm;
                Pl[(m + 1) * (m + 2) / 2 + m] = p1;
                for (int n = m + 2; n <= p; ++n) {
                    const double pn = c_leg.a[n][m] * (ct * p1 - c_leg.b[n][m] * p2);
                    Pl[n * (n + 1) / 2 + m] = pn;
                    p2 = p1;
                    p1 = pn;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kAccPerThread; ++t) {
            const int c = threadIdx.x + t * blockDim.x;
            if (c < nc) {
                int n, m;
                nm_decode(c, n, m);
                const int am = m < 0 ? -m : m;
                const double sg = (m < 0 && (am & 1)) ? -1.0 : 1.0;
                const double sy = m < 0 ? -1.0 : 1.0;
                double2 a = acc[t];
                for (int j = 0; j < nch; ++j) {
                    const double2 qr = s_qrad[j * P1 + n];
                    const double pl = sg * s_leg[j * tri + n * (n + 1) / 2 + am];
                    const double2 ph = s_ph[j * P1 + am];
                    a = cmadd(a, qr, make_double2(pl * ph.x, pl * sy * ph.y));
                }
                acc[t] = a;
            }
        }
        __syncthreads();
    }
    double2* out = M + bx.exp_off[b];
#pragma unroll
    for (int t = 0; t < kAccPerThread; ++t) {
        const int c = threadIdx.x + t * blockDim.x;
        if (c < nc) out[c] = acc[t];
    }
}

// ---------------------------------------------------------------------------------
// L2P: one workgroup (one wave) per leaf, one thread per target point.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void l2p_kernel(const int* leaves, DevBoxes bx, DevLevels lv, const double* X,
                                                 const double* Y, const double* Z, const double2* Lexp, double k,
                                                 double2* far) {
    extern __shared__ double smem[];
    const int b = leaves[blockIdx.x];
    const int L = bx.level[b];
    const int p = lv.p[L];
    const double R = lv.R[L];
    const int P1 = p + 1, nc = P1 * P1;
    double2* s_L = (double2*)smem;
    double* s_rad = (double*)(s_L + nc);
    const double2* src = Lexp + bx.exp_off[b];
    for (int c = threadIdx.x; c < nc; c += blockDim.x) s_L[c] = src[c];
    __syncthreads();
    const double cx = bx.centre[3 * b], cy = bx.centre[3 * b + 1], cz = bx.centre[3 * b + 2];
    const int ps = bx.pstart[b], cnt = bx.count[b];
    double* rad = s_rad + threadIdx.x * P1;
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
        jhat_dev(k * r, p, rad, 1, k * R);
        const double rr = r / R;
        double pw = 1.0;
        for (int n = 0; n <= p; ++n) {
            rad[n] *= pw;
            pw *= rr;
        }
        double2 phi = make_double2(0.0, 0.0);
        double pmm = 0.28209479177387814;
        double2 eim = make_double2(1.0, 0.0);
        for (int m = 0; m <= p; ++m) {
            if (m > 0) {
                pmm *= -c_leg.d[m] * st;
                eim = cmul(eim, eia);
            }
            const double sgn = (m & 1) ? -1.0 : 1.0;
            double p2 = 0.0, p1 = 0.0;
            for (int n = m; n <= p; ++n) {
                double pn;
                if (n == m)
                    pn = pmm;
                else if (n == m + 1)
                    pn = c_leg.e[m] * ct * pmm;
                else
                    pn = c_leg.a[n][m] * (ct * p1 - c_leg.b[n][m] * p2);
                p2 = p1;
                p1 = pn;
                const double g = rad[n] * pn;
                const double2 lp = s_L[n * n + n + m];
                double2 term = cmul(lp, eim);
                if (m > 0) {
                    const double2 lm = s_L[n * n + n - m];
                    const double2 t2 = cmul(lm, conjd(eim));
                    term.x += sgn * t2.x;
                    term.y += sgn * t2.y;
                }
                phi.x = fma(g, term.x, phi.x);
                phi.y = fma(g, term.y, phi.y);
            }
        }
        far[i] = phi;
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
        jhat_dev(k * r, p + 1, rad, 1, k * R);
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
                pmm *= -c_leg.d[m] * st;
                eim = cmul(eim, eia);
            }
            const double sgn = (m & 1) ? -1.0 : 1.0;
            double p2 = 0.0, p1 = 0.0;
            for (int n = m; n <= p + 1; ++n) {
                double pn;
                if (n == m)
                    pn = pmm;
                else if (n == m + 1)
                    pn = c_leg.e[m] * ct * pmm;
                else
                    pn = c_leg.a[n][m] * (ct * p1 - c_leg.b[n][m] * p2);
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
// coordinate differences and r^2 are FP64 (exact to rounding for close pairs), the
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

template <bool PH, bool LOWK, int R>
__device__ __forceinline__ void p2p_mixed_body(const P2PArgs& a, const int leaf, P2PStage& st) {
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
        h.d[m] = m > 0 ? std::sqrt((2.0 * m + 1.0) / (2.0 * m)) : 0.0;
        h.e[m] = std::sqrt(2.0 * m + 3.0);
        for (int n = 0; n <= kMaxP; ++n) {
            if (n >= m + 2) {
                h.a[n][m] = std::sqrt((4.0 * n * n - 1.0) / ((double)n * n - (double)m * m));
                h.b[n][m] = std::sqrt(((n - 1.0) * (n - 1.0) - (double)m * m) / (4.0 * (n - 1.0) * (n - 1.0) - 1.0));
            }
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

void launch_p2m(const int* leaves, int n_items, int chunk, DevBoxes bx, DevLevels lv, const double* X,
                const double* Y, const double* Z, const double2* q, double k, int pmax, double2* M, hipStream_t s) {
    if (n_items <= 0) return;
    upload_legendre_coefficients();
    const size_t shm = (size_t)chunk * p2m_bytes_per_point(pmax);
    hipLaunchKernelGGL(p2m_kernel, dim3(n_items), dim3(256), shm, s, leaves, chunk, bx, lv, X, Y, Z, q, k, M);
    check_launch("p2m");
}

void launch_l2p(const int* leaves, int n_items, DevBoxes bx, DevLevels lv, const double* X, const double* Y,
                const double* Z, const double2* L, double k, int pmax, double2* far, hipStream_t s) {
    if (n_items <= 0) return;
    upload_legendre_coefficients();
    const int P1 = pmax + 1;
    const size_t shm = (size_t)P1 * P1 * sizeof(double2) + 64 * P1 * sizeof(double);
    hipLaunchKernelGGL(l2p_kernel, dim3(n_items), dim3(64), shm, s, leaves, bx, lv, X, Y, Z, L, k, far);
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
