/* Oracle (test infrastructure only): the evaluate phases of oracle/kifmm.py's restatement
 * of the reference's leveled kernel-independent FMM, in C with OpenMP, so that the
 * reference's algorithm can be timed on host cores (setup excluded, as in the metric).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
 *
 * One call evaluates one component (kernel cos(kr)/r or sin(kr)/r, HFMM.py:12-52) for
 * real charges in sorted point order, with the tables oracle/kifmm.py builds:
 *   P2M      jaxfmm/kifmm/phases.py:79-125  (check potentials of leaf points, decode)
 *   M2M      phases.py:128-173              (raw child->parent kernels, parent decode)
 *   M2L      phases.py:176-237              (raw kernels per displacement, one decode)
 *   P2L (X)  phases.py:1709-1770
 *   L2L      phases.py:240-280
 *   L2P      phases.py:1533-1582
 *   M2P (W)  phases.py:1064-1140
 *   near     jaxfmm/near.py:48-312          (U + self pairs, r = 0 excluded)
 * decode(X) = ((X U) * d) Vt (phases.py:21-34).  The M2L runs one displacement class at a
 * time: within a class a target box has at most one source, so the pairs of a class are
 * evaluated in parallel with no write conflicts and the class's matrix stays in cache. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int n, nb, ns, depth, trig;          /* trig 0: cos, 1: sin */
    double k;
    const double *x, *q;                 /* sorted points (n x 3), real charges (n) */
    const int *level, *pstart, *count, *octant;
    const unsigned char* is_leaf;
    const double *centre, *h;            /* box centres (nb x 3), half-sides (nb) */
    const double *eq, *ck;               /* unit equivalent / check surfaces (ns x 3) */
    /* per level l (0..depth), ns x ns row-major unless stated */
    const double *fup_u, *fup_d, *fup_vt;   /* F_up factors: U, d (ns), Vt */
    const double *fdn_u, *fdn_d, *fdn_vt;   /* F_dn factors */
    const double *m2m, *l2l;                /* [l][8][ns][ns] */
    const double *m2l;                      /* [l][343][ns][ns], displacement (dx+3)*49+(dy+3)*7+(dz+3) */
    /* children CSR (parents of every level), V classes, U / W / X lists CSR by target */
    const int *child_ptr, *child;
    int n_vclass;
    const int *vclass_ptr, *vclass_key, *v_tgt, *v_src;   /* key = level * 343 + displacement */
    const int *u_ptr, *u_src, *w_ptr, *w_src, *x_ptr, *x_src;
} kifmm_args;

static inline double kern(const kifmm_args* a, double r) {
    return (a->trig ? sin(a->k * r) : cos(a->k * r)) / r;
}

/* out = ((in U) * d) Vt */
static void decode(int ns, const double* in, const double* U, const double* d, const double* Vt, double* tmp,
                   double* out) {
    for (int j = 0; j < ns; ++j) tmp[j] = 0.0;
    for (int i = 0; i < ns; ++i) {
        const double v = in[i];
        const double* u = U + (size_t)i * ns;
        for (int j = 0; j < ns; ++j) tmp[j] += v * u[j];
    }
    for (int j = 0; j < ns; ++j) tmp[j] *= d[j];
    for (int m = 0; m < ns; ++m) out[m] = 0.0;
    for (int j = 0; j < ns; ++j) {
        const double v = tmp[j];
        const double* w = Vt + (size_t)j * ns;
        for (int m = 0; m < ns; ++m) out[m] += v * w[m];
    }
}

/* y += K x, K ns x ns row-major */
static void matvec_add(int ns, const double* K, const double* x, double* y) {
    for (int i = 0; i < ns; ++i) {
        const double* r = K + (size_t)i * ns;
        double s = 0.0;
        for (int j = 0; j < ns; ++j) s += r[j] * x[j];
        y[i] += s;
    }
}

/* potential at surface points (unit surface * h + c) from the points of box b */
static void surface_from_points(const kifmm_args* a, const double* surf, double hh, const double* c, int b,
                                double* out) {
    const int ns = a->ns, p0 = a->pstart[b], np = a->count[b];
    for (int i = 0; i < ns; ++i) {
        const double sx = surf[3 * i] * hh + c[0], sy = surf[3 * i + 1] * hh + c[1], sz = surf[3 * i + 2] * hh + c[2];
        double s = 0.0;
        for (int j = p0; j < p0 + np; ++j) {
            const double dx = sx - a->x[3 * j], dy = sy - a->x[3 * j + 1], dz = sz - a->x[3 * j + 2];
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (r > 0.0) s += kern(a, r) * a->q[j];
        }
        out[i] = s;
    }
}

/* potential at the points of box t from densities on the surface of box s */
static void points_from_surface(const kifmm_args* a, const double* surf, int s, const double* dens, int t,
                                double* pot) {
    const int ns = a->ns, p0 = a->pstart[t], np = a->count[t];
    const double hh = a->h[s], *c = a->centre + 3 * s;
    for (int i = p0; i < p0 + np; ++i) {
        double acc = 0.0;
        for (int j = 0; j < ns; ++j) {
            const double dx = a->x[3 * i] - (surf[3 * j] * hh + c[0]);
            const double dy = a->x[3 * i + 1] - (surf[3 * j + 1] * hh + c[1]);
            const double dz = a->x[3 * i + 2] - (surf[3 * j + 2] * hh + c[2]);
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (r > 0.0) acc += kern(a, r) * dens[j];
        }
        pot[i] += acc;
    }
}

#define NS2 ((size_t)ns * ns)

int oracle_kifmm_eval(const kifmm_args* a, double* pot /* n, sorted order */) {
    const int ns = a->ns, nb = a->nb, md = a->depth;
    double* up = (double*)calloc((size_t)nb * ns, sizeof(double));
    double* chk = (double*)calloc((size_t)nb * ns, sizeof(double));
    double* dn = (double*)calloc((size_t)nb * ns, sizeof(double));
    if (!up || !chk || !dn) {
        free(up), free(chk), free(dn);
        return -1;
    }
    memset(pot, 0, sizeof(double) * (size_t)a->n);
    /* P2M */
#pragma omp parallel
    {
        double* t1 = (double*)malloc(sizeof(double) * ns);
        double* t2 = (double*)malloc(sizeof(double) * ns);
#pragma omp for schedule(dynamic, 4)
        for (int b = 0; b < nb; ++b) {
            if (!a->is_leaf[b]) continue;
            const int l = a->level[b];
            surface_from_points(a, a->ck, a->h[b], a->centre + 3 * b, b, t1);
            decode(ns, t1, a->fup_u + l * NS2, a->fup_d + (size_t)l * ns, a->fup_vt + l * NS2, t2, up + (size_t)b * ns);
        }
        free(t1);
        free(t2);
    }
    /* M2M: parents of levels md-1 .. 1 */
    for (int pl = md - 1; pl >= 1; --pl) {
#pragma omp parallel
        {
            double* t1 = (double*)malloc(sizeof(double) * ns);
            double* t2 = (double*)malloc(sizeof(double) * ns);
            double* t3 = (double*)malloc(sizeof(double) * ns);
#pragma omp for schedule(dynamic, 4)
            for (int p = 0; p < nb; ++p) {
                if (a->level[p] != pl || a->child_ptr[p] == a->child_ptr[p + 1]) continue;
                for (int i = 0; i < ns; ++i) t1[i] = 0.0;
                for (int e = a->child_ptr[p]; e < a->child_ptr[p + 1]; ++e) {
                    const int c = a->child[e], cl = a->level[c];
                    matvec_add(ns, a->m2m + ((size_t)cl * 8 + a->octant[c]) * NS2, up + (size_t)c * ns, t1);
                }
                decode(ns, t1, a->fup_u + pl * NS2, a->fup_d + (size_t)pl * ns, a->fup_vt + pl * NS2, t2, t3);
                for (int i = 0; i < ns; ++i) up[(size_t)p * ns + i] += t3[i];
            }
            free(t1);
            free(t2);
            free(t3);
        }
    }
    /* M2L by displacement class */
    for (int c = 0; c < a->n_vclass; ++c) {
        const double* K = a->m2l + (size_t)a->vclass_key[c] * NS2;
#pragma omp parallel for schedule(static)
        for (int e = a->vclass_ptr[c]; e < a->vclass_ptr[c + 1]; ++e)
            matvec_add(ns, K, up + (size_t)a->v_src[e] * ns, chk + (size_t)a->v_tgt[e] * ns);
    }
#pragma omp parallel
    {
        double* t1 = (double*)malloc(sizeof(double) * ns);
        double* t2 = (double*)malloc(sizeof(double) * ns);
        double* t3 = (double*)malloc(sizeof(double) * ns);
#pragma omp for schedule(dynamic, 4)
        for (int b = 0; b < nb; ++b) {
            const int l = a->level[b];
            decode(ns, chk + (size_t)b * ns, a->fdn_u + l * NS2, a->fdn_d + (size_t)l * ns, a->fdn_vt + l * NS2, t2,
                   dn + (size_t)b * ns);
            /* P2L: X sources onto the downward check surface (radius R_EQUIV) */
            for (int e = a->x_ptr[b]; e < a->x_ptr[b + 1]; ++e) {
                surface_from_points(a, a->eq, a->h[b], a->centre + 3 * b, a->x_src[e], t1);
                decode(ns, t1, a->fdn_u + l * NS2, a->fdn_d + (size_t)l * ns, a->fdn_vt + l * NS2, t2, t3);
                for (int i = 0; i < ns; ++i) dn[(size_t)b * ns + i] += t3[i];
            }
        }
        free(t1);
        free(t2);
        free(t3);
    }
    /* L2L: children levels 2 .. md */
    for (int cl = 2; cl <= md; ++cl) {
#pragma omp parallel
        {
            double* t1 = (double*)malloc(sizeof(double) * ns);
            double* t2 = (double*)malloc(sizeof(double) * ns);
            double* t3 = (double*)malloc(sizeof(double) * ns);
#pragma omp for schedule(dynamic, 4)
            for (int p = 0; p < nb; ++p) {
                if (a->level[p] != cl - 1) continue;
                for (int e = a->child_ptr[p]; e < a->child_ptr[p + 1]; ++e) {
                    const int c = a->child[e];
                    for (int i = 0; i < ns; ++i) t1[i] = 0.0;
                    matvec_add(ns, a->l2l + ((size_t)cl * 8 + a->octant[c]) * NS2, dn + (size_t)p * ns, t1);
                    decode(ns, t1, a->fdn_u + cl * NS2, a->fdn_d + (size_t)cl * ns, a->fdn_vt + cl * NS2, t2, t3);
                    for (int i = 0; i < ns; ++i) dn[(size_t)c * ns + i] += t3[i];
                }
            }
            free(t1);
            free(t2);
            free(t3);
        }
    }
    /* L2P (downward equivalent surface R_CHECK), M2P (W, upward equivalent R_EQUIV), near */
#pragma omp parallel for schedule(dynamic, 4)
    for (int t = 0; t < nb; ++t) {
        if (!a->is_leaf[t]) continue;
        points_from_surface(a, a->ck, t, dn + (size_t)t * ns, t, pot);
        for (int e = a->w_ptr[t]; e < a->w_ptr[t + 1]; ++e)
            points_from_surface(a, a->eq, a->w_src[e], up + (size_t)a->w_src[e] * ns, t, pot);
        const int p0 = a->pstart[t], np = a->count[t];
        for (int e = a->u_ptr[t] - 1; e < a->u_ptr[t + 1]; ++e) {
            const int s = e < a->u_ptr[t] ? t : a->u_src[e];   /* the leaf itself first */
            const int s0 = a->pstart[s], sn = a->count[s];
            for (int i = p0; i < p0 + np; ++i) {
                double acc = 0.0;
                for (int j = s0; j < s0 + sn; ++j) {
                    const double dx = a->x[3 * i] - a->x[3 * j], dy = a->x[3 * i + 1] - a->x[3 * j + 1];
                    const double dz = a->x[3 * i + 2] - a->x[3 * j + 2];
                    const double r = sqrt(dx * dx + dy * dy + dz * dz);
                    if (r > 0.0) acc += kern(a, r) * a->q[j];
                }
                pot[i] += acc;
            }
        }
    }
    free(up);
    free(chk);
    free(dn);
    return 0;
}
