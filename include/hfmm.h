/* libhfmm: MI355X-native 3-D Helmholtz FMM evaluator, C ABI.
 *
 * Kernel: phi(x_t) = sum_{s, x_s != x_t} q_s exp(i k |x_t - x_s|) / |x_t - x_s|
 * (no 1/(4 pi) factor, self term excluded), as in the reference's HFMM.py:12-52
 * and test.py:11-18.
 *
 * The reference exposes this path as Python functions, not as an FFI; the
 * entry points below are what its `setup` / `evaluate` pair would bind to:
 *   hfmm_plan_create   replaces HFMM.setup            (HFMM.py:55-80)
 *   hfmm_evaluate      replaces HFMM.evaluate         (HFMM.py:83-91)
 *   hfmm_plan_destroy  releases what setup's device_put held (jaxfmm/kifmm/setup.py:191)
 *   hfmm_direct        replaces the FP64 direct sum   (test.py:11-18), on the GPU
 * All calls return 0 on success or a negative status; no exception crosses the ABI.
 * hfmm_last_error() returns the message of the calling thread's last failure.
 * Plans are not thread-safe; distinct plans may be used from distinct threads.
 */
#ifndef HFMM_H
#define HFMM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define HFMM_OK 0
#define HFMM_ERR_ARG (-2)      /* invalid argument: the Python layer raises ValueError */
#define HFMM_ERR_DEVICE (-3)   /* HIP runtime failure or no gfx950 device */
#define HFMM_ERR_MEMORY (-4)   /* device allocation failed */
#define HFMM_ERR_INTERNAL (-9)

/* precision of the evaluate path */
#define HFMM_PREC_FP64 0       /* FP64 everywhere (the plain FMM) */
#define HFMM_PREC_MIXED 1      /* near field: coordinate differences and r^2 in FP32 at the
                                  leaves whose estimated error from them is within
                                  hfmm_opts.p2p_coord32_tol, FP64 elsewhere, weights e^{ikr}/r in
                                  FP32 (the phase k r in FP64 for leaves whose pairs reach
                                  k r > 8), 64-source sums in FP32 added in FP64; M2L in
                                  FP32 on matrix cores (at the levels hfmm_opts.m2l_fp64_levels
                                  leaves in FP32); M2M and L2L in FP32 on matrix cores where the
                                  M2L of every level of the chain is FP32 (hfmm_plan_info says
                                  which), else FP64; P2M/L2P and accumulation in FP64 */

/* pointer kinds for q / out */
#define HFMM_HOST 0
#define HFMM_DEVICE 1

/* charge / output element types */
#define HFMM_F64 0             /* real float64 charges */
#define HFMM_C128 1            /* complex128, interleaved (re, im) */
#define HFMM_F32 2             /* real float32 charges (widened to FP64 on the GPU) */
#define HFMM_C64 3             /* complex64, interleaved (re, im) */

#define HFMM_MAX_LEVELS 24

typedef struct hfmm_opts {
    int p;                          /* truncation degree at every level (>= 2) */
    int p_level[HFMM_MAX_LEVELS];   /* per-level degree; 0 = use p */
    int n_max;                      /* leaf threshold N_max (>= 1) */
    int max_depth;                  /* <= 0: automatic depth (reference rule) */
    int precision;                  /* HFMM_PREC_* */
    int use_graph;                  /* 1: replay evaluate as a captured hipGraph */
    int m2l_kernel;                 /* 0: matrix-core block kernel where it fits, chosen shape;
                                       1: generic kernel; 8, 16, 32, 64: block kernel with that
                                       many pairs per operator application (target boxes x
                                       directions of equal theta and |d|) where it fits */
    int tree_device;                /* 1: build the octree and lists on the GPU (identical result);
                                       0: on the host */
    int m2l_fp64_levels;            /* HFMM_PREC_MIXED only: bit l set keeps the M2L of level l in
                                       FP64; -1 (the default): automatic, FP64 at the levels whose
                                       box diagonal D_l has k D_l > HFMM_MIXED_FP32_MAX_KD */
    int p2p_tile;                   /* HFMM_PREC_MIXED near field: targets per lane (1, 2 or 4);
                                       0: the default (2) */
    double p2p_coord32_tol;         /* HFMM_PREC_MIXED near field, leaves with the phase in FP32:
                                       a leaf takes its coordinate differences and r^2 in FP32
                                       where the estimated relative error of its targets'
                                       near field from that rounding (measured at setup) is at
                                       most this; 0: HFMM_P2P_COORD32_TOL; < 0: FP64 everywhere */
} hfmm_opts;

#define HFMM_P2P_COORD32_TOL 3.0517578125e-05   /* 2^-15 */

/* Largest k D_l (D_l = 2 sqrt(3) extent 2^-(l+1), the box diagonal of level l) whose M2L the
 * automatic rule of HFMM_PREC_MIXED runs in FP32.  A 1e-6 result needs about 0.75 k D + 2 > 26
 * degrees above it, beyond the matrix-core kernel's 23, so such levels run the FP64 generic
 * kernel anyway; the rule keeps lower, user-chosen degrees there in FP64 too.  Measured on
 * 2M points at k = 100 (sphere and clusters): FP32 M2L at k D = 22 and 11 adds nothing
 * visible to the 7e-7 result error. */
#define HFMM_MIXED_FP32_MAX_KD 32.0

typedef struct hfmm_plan hfmm_plan;

typedef struct hfmm_plan_info {
    int64_t n_points;
    int depth;
    int n_boxes;
    int n_leaves;
    int maxleaf;
    int64_t n_v, n_u, n_w, n_x;     /* list sizes (U without self pairs) */
    int64_t n_p2p_pairs;            /* point pairs evaluated directly (incl. r = 0 pairs) */
    int p_level[HFMM_MAX_LEVELS];   /* degree evaluated at each level */
    double extent;
    double pmin[3];
    int p_table[HFMM_MAX_LEVELS];   /* degree the level's tables were built for */
    int m2l_fp32_levels;            /* bit l set: the M2L of level l runs in FP32 */
    int n_leaves_phase64;           /* HFMM_PREC_MIXED: leaves whose FP32 near-field weights take
                                       the phase k r in FP64 (pairs reaching k r > 8, where an FP32
                                       phase errs by ~k r 2^-23 rad); these take coordinate
                                       differences and r^2 in FP64 */
    int m2m_fp32_levels;            /* bit l set: the M2M into level l runs FP32 MFMA arithmetic
                                       (FP64 results); mixed precision, where the M2L of every
                                       level from 2 to l+1 is FP32 and the level is not split
                                       over atomic work items */
    int l2l_fp32_levels;            /* bit l set: the L2L out of level l runs FP32 MFMA arithmetic
                                       (FP64 results), where the M2L of every level from 2 to l+1
                                       is FP32 */
    int n_leaves_coord32;           /* HFMM_PREC_MIXED: leaves whose near field takes coordinate
                                       differences and r^2 in FP32 (hfmm_opts.p2p_coord32_tol) */
} hfmm_plan_info;

void hfmm_opts_default(hfmm_opts* o);

int hfmm_plan_create(const double* xyz /* N x 3, host */, int64_t n, double k, const hfmm_opts* o,
                     hfmm_plan** out);
int hfmm_plan_info_get(const hfmm_plan* plan, hfmm_plan_info* info);

/* Tree export (canonical order: level ascending, Morton code ascending).  Each array
 * pointer may be NULL; sizes are n_boxes (per-box arrays) or n_points (perm). */
int hfmm_plan_export_tree(const hfmm_plan* plan, int* level, uint64_t* code, int* count, int* pstart,
                          int* parent, uint8_t* is_leaf, int64_t* perm);
/* List export: which = 'V', 'U', 'W', 'X'; pairs (target, source) as int32 [n][2]. */
int hfmm_plan_export_list(const hfmm_plan* plan, char which, int* pairs);

/* Evaluate at lower degrees than the tables were built for, without a new setup:
 * p_level[l] in [2, table degree of level l] sets level l; 0 keeps the table degree.
 * Coefficients of degree <= p do not depend on the truncation, so the result equals a
 * plan built at those degrees.  No counterpart in the reference (its tables are built
 * per p in setup, jaxfmm/kifmm/operators.py:572-641). */
int hfmm_plan_set_degrees(hfmm_plan* plan, const int* p_level /* [HFMM_MAX_LEVELS] */);

/* phi for all N points in input order.  q: N charges of q_type (HFMM_F64, HFMM_C128,
 * HFMM_F32 or HFMM_C64; real charges give the reference semantics, HFMM.py:83-91);
 * out: N complex128.  q and out are both host or both device pointers (ptr_kind).
 * stream: a hipStream_t or NULL for the plan's own stream; the call returns after the
 * result is complete. */
int hfmm_evaluate(hfmm_plan* plan, const void* q, int q_type, void* out, int ptr_kind, void* stream);

/* The same with the output element type chosen: out_type HFMM_C128 (as hfmm_evaluate)
 * or HFMM_C64 (the FP64 result rounded to complex64 on the GPU).  The reference returns
 * complex128 only (HFMM.py:91); complex64 is an opt-in for memory-bound callers. */
int hfmm_evaluate_typed(hfmm_plan* plan, const void* q, int q_type, void* out, int out_type, int ptr_kind,
                        void* stream);

/* phi and the field E = -grad phi for all N points (input order): phi N complex128,
 * field N x 3 complex128 (x, y, z per point).  Replaces kifmm.evaluate_setup(...,
 * field=True) (jaxfmm/kifmm/setup.py:229-261; near-field weight HFMM.py:29-37) for the
 * Helmholtz kernel, which HFMM.evaluate does not expose.  Leaf degrees up to 39.  Runs
 * eagerly on the stream (not graph-replayed). */
int hfmm_evaluate_field(hfmm_plan* plan, const void* q, int q_type, void* phi, void* field, int ptr_kind,
                        void* stream);

/* Time the device part of evaluate (inputs already resident): runs `reps` evaluates
 * after `warmup`, returns per-evaluate milliseconds from hipEvents (median) and, if
 * phase_ms is non-NULL, per-phase milliseconds averaged over `reps` further
 * instrumented (non-graph) runs with hipEvents between phases on the plan's stream:
 * [0] gather+clear, [1] P2M, [2] M2M, [3] M2L, [4] L2L, [5] L2P, [6] P2P. */
int hfmm_benchmark(hfmm_plan* plan, const void* q_dev, int q_type, void* out_dev, int warmup, int reps,
                   double* median_ms, double* phase_ms);

/* Measurement helper: duration of the M2L launch of every level, averaged over `reps`
 * instrumented (non-graph) evaluates, hipEvents recorded around each launch on the
 * plan's stream.  level_ms[HFMM_MAX_LEVELS]; levels without M2L read 0.  No counterpart
 * in the reference (its evaluate is one traced function). */
int hfmm_benchmark_m2l_levels(hfmm_plan* plan, const void* q_dev, int q_type, void* out_dev, int reps,
                              double* level_ms);

int hfmm_plan_destroy(hfmm_plan* plan);

/* FP64 direct sum on the GPU at selected targets: out[i] = phi(xyz[targets[i]]) from all
 * n sources.  Host pointers; q complex128 interleaved. */
int hfmm_direct(const double* xyz, int64_t n, const double* q, double k, const int64_t* targets,
                int64_t n_targets, double* out);

/* The same sum on device buffers with no device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under hipGraph capture, serving
 * loops, harnesses that forbid hipMalloc).  xyz: n x 3 FP64, q: n complex128, targets:
 * n_targets int64 (an index outside [0, n) gives NaN), out: n_targets complex128, all in
 * device memory; the kernels run on `stream` (a hipStream_t, NULL for the default stream)
 * and the call returns without waiting.  The partial sums live in a fixed device array of
 * HFMM_DIRECT_NOALLOC_MAX_TARGETS complex128 accumulators (1 MiB) per GPU, held until the
 * process ends and cleared by each call: calls that overlap on one GPU race.  Returns
 * HFMM_ERR_ARG if n_targets exceeds the maximum. */
#define HFMM_DIRECT_NOALLOC_MAX_TARGETS 65536
int hfmm_direct_noalloc(const double* xyz, int64_t n, const double* q, double k, const int64_t* targets,
                        int64_t n_targets, double* out, void* stream);

/* Host-only octree and lists (no device use).  Replaces the reference's tree build
 * and list builders: build_fmm_tree_auto + assign_particles (jaxfmm/trees/cubic.py:530-695),
 * build_v_eager / build_ulist_eager / build_wx_eager (jaxfmm/trees/lists.py:818-931).
 * Export functions take the same arguments as the plan variants above. */
typedef struct hfmm_tree hfmm_tree;
int hfmm_tree_create(const double* xyz, int64_t n, int n_max, int max_depth, hfmm_tree** out);
/* The same tree and lists built on the current GPU (Morton sort, coarsening and list
 * construction in HIP kernels); the exported arrays are identical to hfmm_tree_create's. */
int hfmm_tree_create_device(const double* xyz, int64_t n, int n_max, int max_depth, hfmm_tree** out);
int hfmm_tree_info_get(const hfmm_tree* tree, hfmm_plan_info* info);
int hfmm_tree_export(const hfmm_tree* tree, int* level, uint64_t* code, int* count, int* pstart, int* parent,
                     uint8_t* is_leaf, int64_t* perm);
int hfmm_tree_export_list(const hfmm_tree* tree, char which, int* pairs);
int hfmm_tree_destroy(hfmm_tree* tree);

/* Host FP64 operator-table generators, exposed so the tables can be checked without a
 * GPU (no counterpart in the reference, whose Helmholtz path has no multipole code).
 * jhat_n(x) = j_n(x)(2n+1)!!/x^n and hhat_n(x) = h_n(x)x^{n+1}/(2n-1)!!, n = 0..p;
 * rotation: packed E^n(theta), Y_n^m(R_y(theta)u) = sum_m' E[m][m'] Y_n^m'(u);
 * coax: kind 0 M2L, 1 M2M, 2 L2L, scaled coaxial operator along +z (complex, packed by m). */
int hfmm_host_jhat(double x, int p, double* out);
int hfmm_host_hhat(double x, int p, double* out_re, double* out_im);
int hfmm_host_rotation(double theta, int p, double* out);
int hfmm_host_coax(int kind, double k, double t, double rs, double rt, int ps, int pt, double* out);

/* Degree-predicting MLP (learned truncation degree).  No counterpart in the reference:
 * it takes the role of jaxfmm/tune.py's order selection (tune.py:375-516).
 * 6 inputs -> 3 x 64 GELU -> 1 output; BF16-rounded weights, FP32 accumulation; FP32
 * master weights trained with Adam on minibatches (MSE loss).  Host pointers throughout.
 * Parameter layout: W1[64][6], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[64], b4. */
#define HFMM_MLP_INPUTS 6
#define HFMM_MLP_PARAMS 8833
typedef struct hfmm_mlp hfmm_mlp;
int hfmm_mlp_create(const float* params /* HFMM_MLP_PARAMS */, hfmm_mlp** out);
int hfmm_mlp_params(const hfmm_mlp* mlp, float* out /* HFMM_MLP_PARAMS, FP32 master weights */);
int hfmm_mlp_forward(hfmm_mlp* mlp, const float* x /* n x 6 */, int64_t n, float* y /* n */);
/* gradient of mean((f(x) - y)^2) over the n samples with respect to the rounded weights */
int hfmm_mlp_gradient(hfmm_mlp* mlp, const float* x, const float* y, int64_t n, float* grad, float* loss);
int hfmm_mlp_train(hfmm_mlp* mlp, const float* x, const float* y, int64_t n, int epochs, int batch, float lr,
                   int seed, float* loss_per_epoch /* epochs, may be NULL */);
int hfmm_mlp_destroy(hfmm_mlp* mlp);

/* ---------------------------------------------------------------------------------
 * Batches of independent problems (many uneven problems, e.g. 256 problems of
 * 2k-200k points).  No counterpart in the reference, whose setup/evaluate handle one
 * problem; the role of its tuner driver (jaxfmm/tune.py:32-172) is to run many.
 * One plan per problem (opts[i], k[i]).  hfmm_batch_evaluate runs all problems as one
 * hipGraph (captured once per set of pointers and replayed): the upward and downward
 * passes forked over a few streams balanced by point count, the matrix-core M2L of every
 * problem and level and the near field of every problem as persistent launches whose
 * resident workgroups take (problem, item) work from device queues, one per XCD.
 * q[i] / out[i] as in hfmm_evaluate.
 * --------------------------------------------------------------------------------- */
typedef struct hfmm_batch hfmm_batch;
int hfmm_batch_create(int n_problems, const double* const* xyz, const int64_t* n, const double* k,
                      const hfmm_opts* opts, hfmm_batch** out);
/* the plan of problem i (owned by the batch: do not destroy) */
int hfmm_batch_plan(hfmm_batch* batch, int i, hfmm_plan** plan);
int hfmm_batch_evaluate(hfmm_batch* batch, const void* const* q, int q_type, void* const* out, int ptr_kind,
                        void* stream);
/* median milliseconds of one device-resident batch evaluate (hipEvents, graph replays) */
int hfmm_batch_benchmark(hfmm_batch* batch, const void* const* q_dev, int q_type, void* const* out_dev,
                         int warmup, int reps, double* median_ms);
/* mean milliseconds of the four parts of a device-resident batch evaluate, run eagerly
 * with hipEvents between them: phase_ms[0] upward passes (and generic-kernel M2L levels),
 * [1] the persistent matrix-core M2L, [2] downward passes, [3] the persistent near field */
int hfmm_batch_benchmark_phases(hfmm_batch* batch, const void* const* q_dev, int q_type, void* const* out_dev,
                                int reps, double* phase_ms);
int hfmm_batch_destroy(hfmm_batch* batch);

const char* hfmm_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* HFMM_H */
