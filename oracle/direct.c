/* Oracle (test infrastructure only): FP64 direct sum of the Helmholtz kernel.
 *
 * phi(x_t) = sum_{s : |x_t - x_s| > 0} q_s exp(i k |x_t - x_s|) / |x_t - x_s|
 * E(x_t)   = -grad phi = sum_s q_s (x_t - x_s) (1 - i k r) exp(i k r) / r^3
 *            (the reference kernel bundle's field weight, HFMM.py:29-37, in complex form)
 *
 * Restates the reference's ground truth `direct` (test.py:11-18): complex128
 * result, no 1/(4 pi) factor, the r = 0 term excluded.  Only tests/, smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as a checker or as
 * the timed CPU baseline.  OpenMP over targets; each target sums its sources in
 * index order, so the result does not depend on the thread count.
 */
#include <math.h>
#include <stdint.h>

void oracle_direct(const double *xyz, int64_t n_src, const double *q /* complex interleaved */,
                   double k, const double *tgt_xyz, int64_t n_tgt, double *out /* complex */)
{
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < n_tgt; ++t) {
        const double x = tgt_xyz[3 * t], y = tgt_xyz[3 * t + 1], z = tgt_xyz[3 * t + 2];
        double re = 0.0, im = 0.0;
        for (int64_t s = 0; s < n_src; ++s) {
            const double dx = x - xyz[3 * s], dy = y - xyz[3 * s + 1], dz = z - xyz[3 * s + 2];
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (r > 0.0) {
                const double c = cos(k * r) / r, sn = sin(k * r) / r;
                const double qr = q[2 * s], qi = q[2 * s + 1];
                re += c * qr - sn * qi;
                im += c * qi + sn * qr;
            }
        }
        out[2 * t] = re;
        out[2 * t + 1] = im;
    }
}

void oracle_direct_field(const double *xyz, int64_t n_src, const double *q, double k, const double *tgt_xyz,
                         int64_t n_tgt, double *out /* complex, 3 per target */)
{
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < n_tgt; ++t) {
        const double x = tgt_xyz[3 * t], y = tgt_xyz[3 * t + 1], z = tgt_xyz[3 * t + 2];
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int64_t s = 0; s < n_src; ++s) {
            const double d[3] = {x - xyz[3 * s], y - xyz[3 * s + 1], z - xyz[3 * s + 2]};
            const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (r > 0.0) {
                const double c = cos(k * r), sn = sin(k * r), r3 = r * r * r;
                const double fr = (c + k * r * sn) / r3, fi = (sn - k * r * c) / r3;
                const double qr = q[2 * s], qi = q[2 * s + 1];
                const double gr = fr * qr - fi * qi, gi = fr * qi + fi * qr;
                for (int a = 0; a < 3; ++a) {
                    acc[2 * a] += d[a] * gr;
                    acc[2 * a + 1] += d[a] * gi;
                }
            }
        }
        for (int a = 0; a < 6; ++a) out[6 * t + a] = acc[a];
    }
}
