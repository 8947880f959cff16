// Host-side FP64 generation of the operator tables of the rotation-based Helmholtz FMM.
//
// Expansions are scaled per box radius R = sqrt(3) * h (h = half side), with
// sigma_n = (kR)^n / (2n+1)!!:
//   multipole  Mh_n^m = sum_j q_j jhat_n(k rho_j) (rho_j/R)^n conj(Y_n^m(y_j))
//   local      Lh_n^m such that phi(x) = sum Lh_n^m jhat_n(k r)(r/R)^n Y_n^m(x)
// where jhat_n(x) = j_n(x)(2n+1)!!/x^n and hhat_n(x) = h_n(x) x^{n+1}/(2n-1)!! stay
// O(1) as k -> 0, so k = 0 (the 1/r kernel) is the same code path.  Y_n^m are
// orthonormal spherical harmonics with the Condon-Shortley phase.
//
// Translations are rotate -> coaxial translate along +z -> rotate back.  The coaxial
// coefficients come from the Gegenbauer addition theorem,
//   (S|R)^{n'm}_{nm}(t z) = 4 pi sum_l i^{l+n-n'} h_l(kt) sqrt((2l+1)/4pi) G(n',l,n;m),
//   G(n',l,n;m) = int Y_{n'}^m Y_l^0 conj(Y_n^m) dOmega,
// with the Gaunt integrals evaluated exactly by Gauss-Legendre quadrature; (R|R) uses
// j_l instead of h_l.  All scale factors are folded in log space, so no intermediate
// over- or underflows.  The reference has no Helmholtz multipole code (its Helmholtz
// path is a kernel-independent FMM), so these formulas are derived from the textbook
// addition theorems; they are pinned by the tests against direct summation.
#include <cmath>
#include <complex>
#include <vector>

#include "common.h"

namespace hfmm {

using cplx = std::complex<double>;

void jhat_host(double x, int p, double* out) {
    x = std::fabs(x);
    if (x == 0.0) {
        for (int n = 0; n <= p; ++n) out[n] = 1.0;
        return;
    }
    const int N = p + (int)(1.5 * x) + 30;
    std::vector<double> v(N + 2, 0.0);
    v[N + 1] = 0.0;
    v[N] = 1.0;
    const double x2 = x * x;
    for (int n = N; n >= 1; --n) {
        v[n - 1] = v[n] - x2 * v[n + 1] / ((2.0 * n + 1.0) * (2.0 * n + 3.0));
        if (std::fabs(v[n - 1]) > 1e250)
            for (int j = n - 1; j <= N + 1; ++j) v[j] *= 1e-250;
    }
    double s;
    const double e0 = std::sin(x) / x;
    if (x < 1.0) {
        s = e0 / v[0];
    } else {
        const double e1 = 3.0 * (std::sin(x) - x * std::cos(x)) / (x * x * x);
        s = std::fabs(e0) >= std::fabs(e1) ? e0 / v[0] : e1 / v[1];
    }
    for (int n = 0; n <= p; ++n) out[n] = v[n] * s;
}

void hhat_host(double x, int p, double* re, double* im) {
    std::vector<cplx> h(p + 2);
    const cplx e = std::exp(cplx(0.0, x));
    h[0] = cplx(0.0, -1.0) * e;
    h[1] = -(cplx(x, 1.0)) * e;
    for (int n = 1; n < p; ++n) h[n + 1] = h[n] - x * x * h[n - 1] / ((2.0 * n + 1.0) * (2.0 * n - 1.0));
    for (int n = 0; n <= p; ++n) {
        re[n] = h[n].real();
        im[n] = h[n].imag();
    }
}

// Fully normalised associated Legendre functions with the Condon-Shortley phase,
// P[n*(n+1)/2 + m] for 0 <= m <= n <= p.
static void legendre_table(double ct, double st, int p, double* P) {
    auto at = [&](int n, int m) -> double& { return P[n * (n + 1) / 2 + m]; };
    at(0, 0) = 1.0 / std::sqrt(4.0 * M_PI);
    for (int m = 1; m <= p; ++m) at(m, m) = -std::sqrt((2.0 * m + 1.0) / (2.0 * m)) * st * at(m - 1, m - 1);
    for (int m = 0; m < p; ++m) at(m + 1, m) = std::sqrt(2.0 * m + 3.0) * ct * at(m, m);
    for (int m = 0; m <= p; ++m)
        for (int n = m + 2; n <= p; ++n) {
            const double a = std::sqrt((4.0 * n * n - 1.0) / ((double)n * n - (double)m * m));
            const double b = std::sqrt(((n - 1.0) * (n - 1.0) - (double)m * m) / (4.0 * (n - 1.0) * (n - 1.0) - 1.0));
            at(n, m) = a * (ct * at(n - 1, m) - b * at(n - 2, m));
        }
}

// Gauss-Legendre nodes and weights on [-1, 1] by Newton iteration on P_Q.
static void gauss_legendre(int Q, std::vector<double>& x, std::vector<double>& w) {
    x.resize(Q);
    w.resize(Q);
    for (int i = 0; i < Q; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (Q + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int j = 2; j <= Q; ++j) {
                const double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            dp = Q * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = z;
        for (int j = 2; j <= Q; ++j) {
            const double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
            p0 = p1;
            p1 = p2;
        }
        dp = Q * (z * p1 - p0) / (z * z - 1.0);
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

Gaunt make_gaunt(int p) {
    Gaunt G;
    G.p = p;
    const int L = 2 * p, P1 = p + 1, L1 = L + 1;
    G.g.assign((size_t)P1 * P1 * L1 * P1, 0.0);
    const int Q = 2 * p + 2;                 // exact for polynomial degree <= 4p + 3
    std::vector<double> xs, ws;
    gauss_legendre(Q, xs, ws);
    const int tri = (L + 1) * (L + 2) / 2;
    std::vector<double> P((size_t)Q * tri);
    for (int q = 0; q < Q; ++q)
        legendre_table(xs[q], std::sqrt(std::max(0.0, 1.0 - xs[q] * xs[q])), L, &P[(size_t)q * tri]);
    auto Pq = [&](int q, int n, int m) { return P[(size_t)q * tri + n * (n + 1) / 2 + m]; };
    for (int m = 0; m <= p; ++m)
        for (int n2 = m; n2 <= p; ++n2)
            for (int n = m; n <= p; ++n)
                for (int l = std::abs(n - n2); l <= n + n2; l += 2) {
                    double s = 0.0;
                    for (int q = 0; q < Q; ++q) s += ws[q] * Pq(q, n2, m) * Pq(q, l, 0) * Pq(q, n, m);
                    G.g[(((size_t)m * P1 + n2) * L1 + l) * P1 + n] = 2.0 * M_PI * s;
                }
    return G;
}

int coax_size(int pt, int ps) {
    int s = 0;
    for (int m = 0; m <= std::min(pt, ps); ++m) s += (pt + 1 - m) * (ps + 1 - m);
    return s;
}

int coax_offset(int pt, int ps, int m) {
    int s = 0;
    for (int j = 0; j < m; ++j) s += (pt + 1 - j) * (ps + 1 - j);
    return s;
}

static double log_dfact_odd(int n) {   // log(n!!) for odd n >= -1
    if (n <= 0) return 0.0;
    const int k = (n + 1) / 2;         // (2k-1)!! = (2k)! / (2^k k!)
    return std::lgamma(2.0 * k + 1.0) - k * std::log(2.0) - std::lgamma(k + 1.0);
}

static cplx ipow(int e) {
    switch (((e % 4) + 4) % 4) {
        case 0: return cplx(1, 0);
        case 1: return cplx(0, 1);
        case 2: return cplx(-1, 0);
        default: return cplx(0, -1);
    }
}

void coax_table(int kind, const Gaunt& G, double k, double t, double Rs, double Rt, int ps, int pt, double* out) {
    if (G.p < std::max(ps, pt)) fail(-9, "internal: Gaunt table too small");
    const int lmax = ps + pt;
    std::vector<double> rad(lmax + 1), radi(lmax + 1, 0.0);
    const double x = k * t;
    if (kind == 0)
        hhat_host(x, lmax, rad.data(), radi.data());
    else
        jhat_host(x, lmax, rad.data());
    const double logk = k > 0 ? std::log(k) : 0.0;
    int pos = 0;
    for (int m = 0; m <= std::min(pt, ps); ++m)
        for (int n = m; n <= pt; ++n)
            for (int n2 = m; n2 <= ps; ++n2) {
                cplx s(0, 0);
                for (int l = std::abs(n - n2); l <= n + n2; l += 2) {
                    const double g = G.at(m, n2, l, n);
                    if (g == 0.0) continue;
                    double lg;
                    int pw;
                    if (kind == 0) {
                        lg = n * std::log(Rt / t) + n2 * std::log(Rs / t) + log_dfact_odd(2 * l - 1) -
                             log_dfact_odd(2 * n + 1) - log_dfact_odd(2 * n2 + 1);
                        pw = n + n2 - l;
                        if (pw > 0) {
                            if (x == 0.0) continue;
                            lg += pw * std::log(x);
                        }
                    } else if (kind == 1) {
                        lg = n2 * std::log(Rs) - n * std::log(Rt) + l * std::log(t) + log_dfact_odd(2 * n + 1) -
                             log_dfact_odd(2 * n2 + 1) - log_dfact_odd(2 * l + 1);
                        pw = n2 - n + l;
                        if (pw > 0) {
                            if (k == 0.0) continue;
                            lg += pw * logk;
                        }
                    } else {
                        lg = n * std::log(Rt) - n2 * std::log(Rs) + l * std::log(t) + log_dfact_odd(2 * n2 + 1) -
                             log_dfact_odd(2 * n + 1) - log_dfact_odd(2 * l + 1);
                        pw = n - n2 + l;
                        if (pw > 0) {
                            if (k == 0.0) continue;
                            lg += pw * logk;
                        }
                    }
                    const cplx r(rad[l], radi[l]);
                    s += ipow(l + n - n2) * r * (std::exp(lg) * std::sqrt((2.0 * l + 1.0) / (4.0 * M_PI)) * g);
                }
                cplx v = (kind == 0) ? cplx(0.0, 16.0 * M_PI * M_PI / t) * s : 4.0 * M_PI * s;
                out[2 * pos] = v.real();
                out[2 * pos + 1] = v.imag();
                ++pos;
            }
}

// exp(A) for a small real matrix by scaling and squaring of a Taylor polynomial.
static void expm(int d, std::vector<double>& A, std::vector<double>& E) {
    double nrm = 0.0;
    for (int j = 0; j < d; ++j) {
        double c = 0.0;
        for (int i = 0; i < d; ++i) c += std::fabs(A[i * d + j]);
        nrm = std::max(nrm, c);
    }
    int s = 0;
    while (nrm > 0.5) {
        nrm *= 0.5;
        ++s;
    }
    const double sc = std::ldexp(1.0, -s);
    for (auto& a : A) a *= sc;
    E.assign(d * d, 0.0);
    std::vector<double> T(d * d, 0.0), T2(d * d);
    for (int i = 0; i < d; ++i) E[i * d + i] = T[i * d + i] = 1.0;
    for (int j = 1; j <= 24; ++j) {
        for (int i = 0; i < d; ++i)
            for (int c = 0; c < d; ++c) {
                double acc = 0.0;
                for (int q = 0; q < d; ++q) acc += T[i * d + q] * A[q * d + c];
                T2[i * d + c] = acc / j;
            }
        T.swap(T2);
        for (int i = 0; i < d * d; ++i) E[i] += T[i];
    }
    for (int r = 0; r < s; ++r) {
        for (int i = 0; i < d; ++i)
            for (int c = 0; c < d; ++c) {
                double acc = 0.0;
                for (int q = 0; q < d; ++q) acc += E[i * d + q] * E[q * d + c];
                T2[i * d + c] = acc;
            }
        E.swap(T2);
    }
}

void rotation_table(double theta, int p, double* out) {
    for (int n = 0; n <= p; ++n) {
        const int d = 2 * n + 1;
        std::vector<double> A(d * d, 0.0), E;
        for (int m = -n; m < n; ++m) {
            const double c = 0.5 * std::sqrt((double)n * (n + 1) - (double)m * (m + 1));
            // generator G with G[m+1][m] = c, G[m][m+1] = -c; E = exp(-theta G)
            A[(m + n + 1) * d + (m + n)] = -theta * c;
            A[(m + n) * d + (m + n + 1)] = theta * c;
        }
        expm(d, A, E);
        double* o = out + rot_offset(n);
        for (int i = 0; i < d * d; ++i) o[i] = E[i];
    }
}

// Blocks of a rotation (E, or E^T when transpose) in the basis u_0 = x_0,
// u_m = x_m + (-1)^m x_{-m}, v_m = x_m - (-1)^m x_{-m} (m = 1..n), where it is block
// diagonal: Wu (n+1) x (n+1), row-major by (output m, input m); Wv n x n by (m-1, m-1).
// Column j of a block is the image of one basis vector: x = e_0 for u_0, and
// x_j = 1/2, x_{-j} = +-(-1)^j / 2 for u_j / v_j.
void rotation_uv_blocks(const double* E, int n, bool transpose, std::vector<double>& Wu, std::vector<double>& Wv) {
    const int w = 2 * n + 1;
    auto B = [&](int a, int b) { return transpose ? E[(b + n) * w + (a + n)] : E[(a + n) * w + (b + n)]; };
    auto sgn = [](int m) { return (m & 1) ? -1.0 : 1.0; };
    Wu.assign((size_t)(n + 1) * (n + 1), 0.0);
    Wv.assign((size_t)n * n, 0.0);
    std::vector<double> y(w);
    for (int j = 0; j <= n; ++j) {
        for (int a = -n; a <= n; ++a) y[a + n] = j == 0 ? B(a, 0) : 0.5 * (B(a, j) + sgn(j) * B(a, -j));
        Wu[j] = y[n];
        for (int i = 1; i <= n; ++i) Wu[(size_t)i * (n + 1) + j] = y[i + n] + sgn(i) * y[n - i];
    }
    for (int j = 1; j <= n; ++j) {
        for (int a = -n; a <= n; ++a) y[a + n] = 0.5 * (B(a, j) - sgn(j) * B(a, -j));
        for (int i = 1; i <= n; ++i) Wv[(size_t)(i - 1) * n + (j - 1)] = y[i + n] - sgn(i) * y[n - i];
    }
}

}  // namespace hfmm
