// Host build of the near-field weight functions (hfmm0_amd/csrc/p2p_math.h), the same
// source the GPU kernels include, exported over arrays for tests/test_nearfield_math.py.
#include "../../hfmm0_amd/csrc/p2p_math.h"

using namespace hfmm::nearfield;

extern "C" {
void nf_sincospi_half(const float* y, int n, float* s, float* c) {
    for (int i = 0; i < n; ++i) sincospi_half(y[i], s[i], c[i]);
}
void nf_halfturn(const float* u2, int n, float* wr, float* wi) {
    for (int i = 0; i < n; ++i) halfturn_weight(u2[i], wr[i], wi[i]);
}
void nf_halfturn_phase64(const double* u2, int n, float* wr, float* wi) {
    for (int i = 0; i < n; ++i) halfturn_weight_phase64(u2[i], wr[i], wi[i]);
}
void nf_pair_weight(const double* r2, int n, double fl2, int ph, float* wr, float* wi) {
    for (int i = 0; i < n; ++i) {
        if (ph)
            pair_weight<true>(r2[i], fl2, wr[i], wi[i]);
        else
            pair_weight<false>(r2[i], fl2, wr[i], wi[i]);
    }
}
void nf_sincos_halfpi(const float* y, int n, float* s, float* c) {
    for (int i = 0; i < n; ++i) sincos_halfpi(y[i], s[i], c[i]);
}
void nf_radian(const float* r2, int n, float k, float* wr, float* wi) {
    for (int i = 0; i < n; ++i) radian_weight(r2[i], k, wr[i], wi[i]);
}
void nf_radian_phase64(const double* r2, int n, double k, float* wr, float* wi, float* rinv, float* kr) {
    for (int i = 0; i < n; ++i) radian_weight_phase64(r2[i], k, wr[i], wi[i], rinv[i], kr[i]);
}
}
