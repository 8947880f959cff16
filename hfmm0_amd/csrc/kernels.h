// Launchers of the HIP kernels (plain pointers; called from the host plan code).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hfmm {

// Per-box data, device side (canonical box order).
struct DevBoxes {
    const double* centre;      // 3 per box
    const int* level;
    const int* pstart;
    const int* count;
    const int64_t* exp_off;    // offset (in complex coefficients) of the box's expansion
};

// Per-level data, device side.
struct DevLevels {
    const int* p;              // truncation degree per level
    const double* R;           // expansion radius sqrt(3) h per level
};

// Displacement information for one translation direction.
struct DispInfo {
    int theta;                 // index into the rotation tables
    int dist;                  // index of the distance within a level's coaxial tables
};

// Copies the Legendre recurrence coefficients to constant memory (once per process).
void upload_legendre_coefficients();

void launch_fill_zero(double2* buf, int64_t n, hipStream_t s);
void launch_gather_charges(const void* q, int q_type, const int* perm, int64_t n, int conj, double2* q_sorted,
                           hipStream_t s);

// LDS bytes per staged point of the P2M kernel at degree p
inline size_t p2m_bytes_per_point(int p) {
    const size_t P1 = p + 1;
    return 24 * P1 + 32;
}
// points staged per pass of the P2M kernel at the largest leaf degree pmax: the staged
// points and the recurrence coefficients stay within the LDS budget of its workgroup size
int p2m_chunk(int pmax);
// One leaf of the P2M and L2P launches, 64 bytes, read by 16 lanes with one vector load.
struct LeafDesc {
    int box, level, pstart, count;
    int64_t exp_off;           // offset of the leaf's expansion
    int64_t pad0;
    double cx, cy, cz;         // box centre
    double pad1;
};
static_assert(sizeof(LeafDesc) == 64, "LeafDesc is 16 words");

// P2M and L2P run as persistent launches: a grid of resident workgroups, each taking the
// leaves of a contiguous range on its XCD in turn (pmax: the largest leaf degree).
void launch_p2m(const LeafDesc* leaves, int n_items, int chunk, DevLevels lv, const double* X, const double* Y,
                const double* Z, const double2* q, double k, int pmax, double2* M, hipStream_t s);

void launch_l2p(const LeafDesc* leaves, int n_items, DevLevels lv, const double* X, const double* Y, const double* Z,
                const double2* L, double k, int pmax, double2* far, hipStream_t s);

// Generic rotate / coaxial / rotate-back translation over a CSR pair list.
// Target t = targets[blockIdx.x]; sources src[ptr[t_i] .. ptr[t_i+1]) with direction dir[].
// dst[exp_off[t]] (+)= sum over sources of the translated expansion.
struct XlateArgs {
    const int* targets;        // n_targets box ids
    const int* ptr;            // CSR offsets indexed by target ordinal
    const int* src;            // source box ids
    const unsigned short* dir; // direction index per entry
    const DispInfo* dinfo;     // per direction
    const double2* phase;      // per direction, e^{i m phi}, m = 0..pmax
    const double* rot;         // rotation tables, rot_size(pmax) doubles per theta index
    const double2* coax;       // coaxial tables of this launch, indexed by dinfo.dist
    int64_t coax_stride;       // complex entries per coaxial table
    int ps, pt, pmax;          // effective source / target degrees (<= table degrees)
    int tps, tpt;              // degrees the coaxial tables were built for
    int accumulate;            // 1: dst += result, 0: dst = result
    int split;                 // workgroups per target (> 1: contiguous parts of the source
                               // list, added atomically; needs accumulate = 1)
    const double2* srcbuf;
    double2* dstbuf;
    const int64_t* exp_off;
};
void launch_xlate(const XlateArgs& a, int n_targets, hipStream_t s);

// the field E = -grad phi (3 complex values per point, sorted order for far_e, input
// order for field) beside the potential
void launch_l2p_field(const int* leaves, int n_items, DevBoxes bx, DevLevels lv, const double* X, const double* Y,
                      const double* Z, const double2* L, double k, int pmax, double2* far, double2* far_e,
                      hipStream_t s);
// Near field over the leaves in leaf_box order, coordinate differences and r^2 in FP64.
// FP32 (precision != 0): FP32 weights; the first n_fp32 leaves with an FP32 phase k r, the
// rest with the phase in FP64 (their pairs reach k r > kP2PFp32MaxKr).  FP64: every leaf
// in FP64.  Pairs with r^2 <= floor2[leaf] (the coincidence floor of the leaf's pairs,
// in unscaled coordinates) contribute nothing.
constexpr double kP2PFp32MaxKr = 8.0;
// Largest k r over a plan's near pairs for which the mixed near field takes e^{ikr} as
// 1 + ikr (relative error (kr)^2 / 2 <= 5e-9, below FP32 rounding).
constexpr double kP2PLowKr = 1e-4;
void launch_p2p_field(int precision, int n_leaves, int n_fp32, const int* leaf_box, DevBoxes bx,
                      const int* rng_ptr, const int* rng_start, const int* rng_pref, const int* leaf_src_total,
                      const double* floor2, const double* X, const double* Y, const double* Z, const double2* q,
                      double k, const double2* far, const double2* far_e, const int* perm, int out_mode,
                      double2* out, double2* field, hipStream_t s);

struct P2PArgs {
    int leaf0;                    // first leaf of the launch (set by the launcher)
    const int* leaf_box;          // leaf -> box
    DevBoxes bx;
    const int* rng_ptr;           // per leaf: its source point ranges [rng_ptr[l], rng_ptr[l+1])
    const int* rng_start;         // first point of each range
    const int* rng_pref;          // sources of the leaf before each range
    const int* leaf_src_total;    // sources of each leaf
    const double* floor2;         // per leaf: coincidence floor on r^2 (unscaled coordinates)
    const double *X, *Y, *Z;      // sorted points (FP64)
    const double2* q;             // sorted charges (complex FP64)
    double k;                     // |k|
    double scale;                 // coordinate scale of the mixed kernel (set by the launcher)
    double wscale;                // factor of the mixed kernel's sums (set by the launcher)
    const double2* far;           // far field per sorted point
    const int* perm;              // sorted -> input order
    int out_mode;                 // bit 0: conjugate, bit 1: zero imaginary part
    double2* out;                 // input order
    const unsigned char* coord32; // per leaf (mixed form, FP32 phase): 1 = coordinate differences
                                  // and r^2 in FP32 (null: FP64 at every leaf)
};
// lowk: every near pair of the plan has k r <= kP2PLowKr (mixed form only); tile: targets
// per lane of the mixed form (1, 2 or 4)
void launch_p2p(const P2PArgs& a, int precision, int n_leaves, int n_fp32, bool lowk, int tile, hipStream_t s);
// Per leaf l < n (leaf order of a): flag[l] = 1 where FP32 coordinates keep the estimated
// relative near-field error of every target of the leaf at or below tol:
//   2^-21 c_max sqrt(sum_j r_j^-4 / sum_j r_j^-2)  <= tol,
// c_max the largest coordinate magnitude relative to the leaf centre over its targets and
// sources (2^-21 c_max bounds the error of an FP32 coordinate difference), the sums over the
// target's sources beyond the coincidence floor (the rounding error of pair j perturbs its
// weight by about r_j^-2 times that error; with random charges the perturbations add in
// quadrature against a potential of about sqrt(sum r_j^-2)).  FP64 arithmetic; setup only.
void launch_p2p_coord_check(const P2PArgs& a, int n, double tol, unsigned char* flag, hipStream_t s);

void launch_direct(const double* X, const double* Y, const double* Z, const double2* q, int64_t n, double k,
                   const double* tx, const double* ty, const double* tz, int64_t nt, double2* out, hipStream_t s);

// M2L on matrix cores over target blocks (m2l_mfma.hip).  The operator tables are
// stored as MFMA A fragments: for a tile (row tile mt, k group kg of four 4-wide
// k-steps) lane l holds, as four consecutive values j = 0..3,
// A[16 mt + l % 16][16 kg + 4 j + l / 16], zero outside the block.
//
// Rotations act on each degree block in the basis u_m = x_m + (-1)^m x_{-m} (m >= 1),
// u_0 = x_0, v_m = x_m - (-1)^m x_{-m} (m >= 1): the reflection x_m -> (-1)^m x_{-m}
// commutes with the rotation exp(-theta G) (G the tridiagonal generator, whose entries are
// symmetric under m -> -m-1), so E_n splits into an (n+1) x (n+1) block on the u and an
// n x n block on the v.  In the LDS buffer u_m sits in the row of x_m and v_m in the row of
// x_{-m}: the u block of degree n is rows n^2+n .. n^2+2n (m ascending), the v block rows
// n^2 .. n^2+n-1 (m = n down to 1).
constexpr int kM2LBlockMaxP = 23;  // largest degree handled by the block kernel
// FP32 one-tile instances: blocks of at most this many rows (forward rotation) or of at
// most this size squared (coaxial) run on the 4x4x1 MFMA in 4-row tiles.  The fragments
// beyond a block's first are loaded when used; with 8 (two tiles) those loads cost more
// than the MFMA passes saved (level-5 M2L at p = 11: 4.38 against 3.90 ms).
constexpr int kM2LQuadMaxRows = 4;
constexpr int kXcds = 8;            // accelerator dies; workgroups are dealt to them round-robin

// Work queues of a persistent launch (batches of problems): work items [start[x],
// start[x+1]) form queue x, one per XCD, each in spatial order; count[16 x] is the next
// index of queue x (one 64-byte line per queue), cleared before every launch.
struct WorkQueues {
    const int2* work;             // (descriptor, item) pairs
    const int* start;             // kXcds + 1 entries
    int* count;                   // 16 kXcds ints
};
#ifdef __HIP__
// The next work item for a workgroup on XCD `home`: its own queue first, then the others'
// (so that the last items are shared out); -1 when every queue is empty.  Indices only
// grow, so every caller returns after at most one pass over the queues.
__device__ inline int work_dequeue(const WorkQueues& q, int home) {
    for (int t = 0; t < kXcds; ++t) {
        const int x = (home + t) % kXcds;
        const int n = q.start[x + 1] - q.start[x];
        int* c = q.count + 16 * x;
        if (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= n) continue;
        const int i = atomicAdd(c, 1);
        if (i < n) return q.start[x] + i;
    }
    return -1;
}
#endif
void launch_clear_ints(int* p, int n, hipStream_t s);
constexpr int kM2LMaxAccRegs = 128; // VGPRs per lane for the stage-3 accumulators (block-size choice)
__host__ __device__ constexpr int m2l_tiles(int len) { return (len + 15) / 16; }
// stage-3 units per wave the M2L block kernel is instantiated for at table degree p: the
// units (n, row tile of the n+1 orders), sum_{n<=p} m2l_tiles(n+1), dealt to 4 waves
__host__ __device__ constexpr int m2l_block_maxu(int p) {
    return p <= 7 ? 2 : p <= 11 ? 3 : p <= 15 ? 4 : p <= 17 ? 5 : p <= 19 ? 6 : 8;
}
// FP32 one-tile instances (p <= 15) pack the stage-3 units of the degrees n1 < n2 with
// n1 + n2 = min(p_tab, kM2LPackMax) into one 16-row tile: rows 0..7 take the orders of n1,
// rows 8..15 those of n2 (unit code 256 | n1 << 4 | n2; an odd count leaves the middle
// degree alone), which lowers the units per wave to these
constexpr int kM2LPackMax = 7;
__host__ __device__ constexpr int m2l_block_maxu_packed(int p) { return p <= 7 ? 1 : p <= 11 ? 2 : 3; }
// accumulator VGPRs of a (MAXU units per wave, TB targets, element size) configuration:
// two (u and v) 16x16 tiles per unit and direction column tile
__host__ __device__ constexpr int m2l_acc_regs(int maxu, int tb, int elem_bytes) {
    return maxu * 2 * (2 * tb / 16) * 4 * (elem_bytes / 4);
}
// LDS rows of the expansion block: (p+1)^2 expansion rows, a scratch row that receives
// the stores of rows outside a block, a zero row (read for orders beyond the evaluated
// degree), and as many further zero rows as the padded tiles of the last u block read:
// those reads end at row p^2 + p + 16 m2l_tiles(p+1) - 1.  (At p = 11 in FP32 with 16 x 2
// pairs per group: 148 rows, 38.5 KB, four workgroups per CU.)
__host__ __device__ constexpr int m2l_lds_x_rows(int p) {
    return (p + 1) * (p + 1) + 2 > p * p + p + 16 * m2l_tiles(p + 1) ? (p + 1) * (p + 1) + 2
                                                                      : p * p + p + 16 * m2l_tiles(p + 1);
}
// phase table of the M2L block kernel: entries per (lane parity, direction) for the
// orders m = 0 .. 16 KT - 1 (KT = m2l_tiles(p + 1))
__host__ __device__ constexpr int m2l_phase_len(int p) { return 16 * m2l_tiles(p + 1); }
__host__ __device__ constexpr int m2l_lds_x_elems(int p, int ncol) {
    return (m2l_lds_x_rows(p) * (ncol + 1) + 3) & ~3;
}
struct M2LBlockArgs {
    int p;                           // effective source degree (<= p_tab)
    int p_tab;                       // source degree the tables were built for
    int pt;                          // effective target degree (<= pt_tab); = p for M2L
    int pt_tab;                      // target degree the tables were built for
    int overwrite;                   // store instead of add (every block in one item)
    int n_items;
    const int4* items;               // (block, first group, end group, atomic add)
    const int* blk_targets;          // blocks x TB target boxes (-1: empty slot)
    const unsigned short* grp_dir;   // per group: G directions of equal theta and |d| (0xFFFF: none)
    const int* grp_off;              // per group: G x TB source expansion offsets (-1: none)
    const int* units1;               // forward rotation units: n | block << 8 (0: u, 1: v)
    int n_units1;
    const int* units2;               // coaxial units: |m| (orders +|m| and -|m|)
    int n_units2;
    const int* units3;               // back rotation units: n | row tile << 8 (u and v)
    int n_units3;
    const void* rot1;                // per rotation: fragments of the u and v blocks of E_n^T (T)
    const void* rot3;                // per rotation: fragments of the u and v blocks of E_n (T)
    const int* rot_off1;             // offsets of (n, block) [2 n + block] within rot1's rotation
    const int* rot_off1q;            // blocks of at most 4 rows: their 4x4x1 fragments (-1: larger)
    const int* rot_off3;             // the same within rot3's
    const int* rot_off3p;            // packed stage-3 units (n1, n2): [8 n1 + n2] within rot3's
    int rot_stride1, rot_stride3;
    const void* phase;               // e^{i m phi} per direction (T complex)
    int phase_stride;
    const DispInfo* dinfo;
    const void* coax;                // per distance: fragments of T^m, re then im per lane (T)
    const int* coax_off;             // offset of order m's fragments within one distance
    const int* coax_offq;            // blocks of at most 4 x 4: their 4x4x1 fragments (-1: larger)
    int coax_stride;
    const void* M;                   // multipoles (T complex)
    double2* L;                      // local expansions (FP64), accumulated into
    const int64_t* exp_off;
};
size_t m2l_block_lds_bytes(int p, int tb, int g, int precision_bytes);
void prepare_m2l_block(int p, int tb, int g, int fp32);
void launch_m2l_block(const M2LBlockArgs& a, int tb, int g, int fp32, hipStream_t s);
// Persistent form over many translations (batches): the resident grid of the instance for
// (p = larger table degree, tb, g, fp32) with `shm` bytes of LDS, and the launch over the
// device descriptors d_args and the queues q of (descriptor, item) work.
int m2l_block_persistent_grid(int p, int tb, int g, int fp32, size_t shm);
// near field of one problem in a persistent batch launch
constexpr int kP2PBatchTargetsPerLane = 2;
struct P2PBatchArgs {
    P2PArgs a;                    // scale and wscale set (p2p_scaled)
};
// leaf classes of the persistent near field: FP64 plans; mixed plans with k r <= kP2PLowKr;
// mixed, FP32 phase; mixed, FP64 phase
constexpr int kP2PClassFp64 = 0, kP2PClassLowK = 1, kP2PClassPhase32 = 2, kP2PClassPhase64 = 3;
P2PArgs p2p_scaled(P2PArgs a, int precision, bool lowk);
int p2p_persistent_grid(int cls);
void launch_p2p_persistent(int cls, const P2PBatchArgs* d_args, const WorkQueues& q, int grid, hipStream_t s);
void launch_m2l_block_persistent(const M2LBlockArgs* d_args, const WorkQueues& q, int grid, int p, int tb, int g,
                                 int fp32, size_t shm, hipStream_t s);
void launch_to_float(const double2* in, float2* out, int64_t n, hipStream_t s);
// zero the first ncoef coefficients of the expansions of the given boxes
void launch_zero_boxes(const int* boxes, int n, const int64_t* exp_off, int ncoef, double2* E, hipStream_t s);
// the result rounded to complex64 (output opt-in)
void launch_to_c64(const double2* in, float2* out, int64_t n, hipStream_t s);

}  // namespace hfmm
