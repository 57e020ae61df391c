// si_device.hpp -- device-resident data structures and kernel launchers.
//
// HBM layout (one GPU):
//   Laplacian L = I - A,  A = D^-1/2 W D^-1/2 with W's diagonal dropped
//     rowptr  int32[n+1]   (nnz < 2^31)
//     col     int32[nnz]   columns, ascending within a row
//     a32     f32[nnz]     A's entries (positive), used by the FP32 bulk sweeps
//     l64     f64[nnz]     L's off-diagonal entries, -isd_i * w_ij * isd_j in FP64,
//                          the exact values of the reference's f64 Laplacian
//     deg, sqrt_deg, isd  f64[n]
//   The unit diagonal is never stored: every SpMM adds x_i in its epilogue.
//   Dense blocks are row-major [n x ld] (ld >= width), FP32 or FP64.
#pragma once

#include <utility>
#include <vector>

#include "si_common.hpp"

namespace si {

constexpr int kLongRow = 256;  // rows above this length get a workgroup of their own

struct DevLaplacian {
    int64_t n = 0;
    int64_t nnz = 0;
    int32_t* rowptr = nullptr;
    int32_t* col = nullptr;
    float* a32 = nullptr;
    double* l64 = nullptr;
    double* deg = nullptr;
    double* sqrt_deg = nullptr;
    double* isd = nullptr;
    int32_t* long_rows = nullptr;  // indices of rows with > kLongRow entries
    int64_t n_long = 0;
    int64_t max_row_len = 0;
    int symmetric = 0;       // W bitwise symmetric with sorted rows
    int degree_mode = 0;     // 0 PythonCompat, 1 RustNative
    int degree_exact = 1;    // PythonCompat degrees reproduce the f32 column-sum order
    int32_t* perm = nullptr;   // perm[new] = old node id (nullptr: identity order)
    int32_t* iperm = nullptr;  // iperm[old] = new node id
    int order = 0;             // storage order: 0 input, 1 reverse Cuthill-McKee, 2 clustering
    // connected components, when the device locality order computed them
    // (n_comp = 0: not computed): comp[old node id] = label, numbered by
    // smallest node
    int64_t n_comp = 0;
    int32_t* comp = nullptr;
    // FP32 wide-SpMM shapes for 16 and 32 columns (index into the candidate
    // list of k_spmm.hip; 0 = default), set by spmm_autotune; tuned holds the
    // widths already tuned (kTuned16 | kTuned32)
    int shape16 = 0;
    int shape32 = 0;
    int tuned = 0;
};

// ----------------------------------------------------------------- laplacian
// Builds L on the device from a CSR graph already resident on the device
// (int64 row pointers, int32 columns, f32 weights).  Throws si::Error on
// malformed input.
// reorder = true renumbers the nodes for gather locality (degrees are computed
// in the input order first, so PythonCompat stays bit-exact): a connected
// symmetric graph of at least kClusterOrderMin nodes and average degree
// kClusterOrderMinDegree takes the clustering order (unless prefer_rcm), other
// symmetric graphs reverse Cuthill-McKee on
// the device, an asymmetric one reverse Cuthill-McKee on the host, where h_*
// optionally give the same pattern.
void laplacian_build_device(DevLaplacian& L, int64_t n, const int64_t* d_indptr,
                            const int32_t* d_indices, const float* d_w, int64_t nnz, int mode,
                            hipStream_t s, bool reorder = false,
                            const int64_t* h_indptr = nullptr, const int32_t* h_indices = nullptr,
                            bool prefer_rcm = false);
// Device Laplacian from host FP64 off-diagonal values (CSR without the
// diagonal; unit diagonal implied) and FP64 degrees; input order is kept.
void laplacian_from_values_device(DevLaplacian& L, int64_t n, const int64_t* h_ip,
                                  const int32_t* h_ix, const double* h_l, const double* h_deg,
                                  hipStream_t s);
void laplacian_free(DevLaplacian& L);
// perm[new] = old of a symmetric CSR pattern (reverse Cuthill-McKee) on up to
// `threads` host threads (0: as the hardware allows, at most 16); the order
// does not depend on the thread count.
std::vector<int32_t> rcm_order(int64_t n, const int64_t* ip, const int32_t* ix, int threads = 0);
// The permuted CSR P W P^T on the device: new row i is original row perm[i]
// with its columns renumbered through iperm and sorted (stable, so equal
// columns keep their input order).  rp2 has n + 1 entries; c2, w2 nnz.
void permute_csr_device(int64_t n, int64_t nnz, const int64_t* rp, const int32_t* col,
                        const float* w, const int32_t* perm, const int32_t* iperm, int32_t* rp2,
                        int32_t* c2, float* w2, hipStream_t s);
// Connected components of a symmetric CSR pattern on the device: lab[v] =
// component id, numbered by smallest node (the labels of a BFS sweep in node
// order); root[c] (optional, n entries) = the smallest node of component c.
// Returns the number of components.
int64_t components_device(int64_t n, const int64_t* rp, const int32_t* ix, int32_t* lab,
                          int32_t* root, hipStream_t s);
// The order of rcm_order for a symmetric pattern, computed on the device:
// perm[new] = old and iperm[old] = new (device arrays of n).
// lab (optional, device, n entries) receives the component labels; returns
// the number of components.
int64_t rcm_order_device(int64_t n, const int64_t* rp, const int32_t* ix, int32_t* perm,
                         int32_t* iperm, hipStream_t s, int32_t* lab = nullptr);
// The same order from known components: lab (n) and the nc smallest nodes in
// start (device, overwritten).
void rcm_order_labeled(int64_t n, const int64_t* rp, const int32_t* ix, const int32_t* lab,
                       int32_t* start, int64_t nc, int32_t* perm, int32_t* iperm, hipStream_t s);
// The clustering order of a connected graph (rows grouped by k-means clusters
// of a 16-column random-walk embedding, node order inside each cluster); L is
// the operator in the input node order with its values built; n >= 256.
void cluster_order_device(const DevLaplacian& L, int32_t* perm, int32_t* iperm, hipStream_t s);
// Connected graphs of at least this many nodes and this average degree (stored
// entries per node) take the clustering order.
constexpr int64_t kClusterOrderMin = 65536;
constexpr int64_t kClusterOrderMinDegree = 8;
// dst[i, :k] = src[idx[i], :k] (row gather); idx == nullptr copies.
void gather_rows_f64(const double* src, int lds, double* dst, int ldd, int64_t n, int k,
                     const int32_t* idx, hipStream_t s);
void gather_rows_f32(const float* src, int lds, float* dst, int ldd, int64_t n, int k,
                     const int32_t* idx, hipStream_t s);

// ----------------------------------------------------------------- SpMM
// out = s_l * (L x) + s_x * x + s_z * z   for row-major blocks of width k.
// x, z, out have leading dimension ld (>= k).  z may be null when s_z == 0.
// With gate set (a device int), the launch does nothing when gate_step > *gate
// (read when the kernel starts: work of a device-planned round).
void spmm_f32(const DevLaplacian& L, const float* x, float* out, int k, int ld, float s_l,
              float s_x, float s_z, const float* z, hipStream_t s,
              const int32_t* gate = nullptr, int gate_step = 0);
void spmm_f64(const DevLaplacian& L, const double* x, double* out, int k, int ld, double s_l,
              double s_x, double s_z, const double* z, hipStream_t s,
              const int32_t* gate = nullptr, int gate_step = 0);
// Segmented form (a block-diagonal L whose diagonal blocks are row segments):
// row i uses (s_l, s_x, s_z) = coef[3 * rseg[i] + {0, 1, 2}] (device arrays);
// a segment with s_l == 0 copies s_x * x + s_z * z without gathering.
void spmm_seg_f32(const DevLaplacian& L, const float* x, float* out, int k, int ld,
                  const float* z, const int32_t* rseg, const float* coef, hipStream_t s,
                  const int32_t* gate = nullptr, int gate_step = 0);
void spmm_seg_f64(const DevLaplacian& L, const double* x, double* out, int k, int ld,
                  const double* z, const int32_t* rseg, const double* coef, hipStream_t s,
                  const int32_t* gate = nullptr, int gate_step = 0);
// Step `step` of a device-planned segmented filter (one launch): with
// g = *gate the step runs only while step <= g; it reads the previous step's
// output (x0 for step 1) and the one before (x0 for step 2, none for step 1)
// from the three-buffer ring and writes ring[(g - step) mod 3], so the round's
// last step always leaves its result in ring[0].  coef is this step's
// [(nseg + 1) x 3] scalar row.  resident: launch only the resident workgroups
// (an idle launch then costs a few microseconds; a busy one is a little
// slower than one workgroup per row block).  k must be a single-launch width.
void spmm_seg_ring_f32(const DevLaplacian& L, const float* x0, float* const ring[3], int k, int ld,
                       const int32_t* rseg, const float* coef, const int32_t* gate, int step,
                       bool resident, hipStream_t s);
void spmm_seg_ring_f64(const DevLaplacian& L, const double* x0, double* const ring[3], int k,
                       int ld, const int32_t* rseg, const double* coef, const int32_t* gate,
                       int step, bool resident, hipStream_t s);
bool spmm_width_supported(int k);
// y = A x for a general CSR matrix (FP64 values, any stored entries) with the
// products rounded before each sum in storage order (the scalar spmv_csr's
// arithmetic, bit for bit); all arrays on the device.
void csr_spmv_exact_f64(int64_t n, const int64_t* ip, const int32_t* ix, const double* data,
                        const double* x, double* y, hipStream_t s);
// true while an SpmmProfile is collecting (SpMM calls then record events)
bool spmm_profiling();
// Pick the FP32 16- and 32-column SpMM shapes for this graph by timing every
// candidate on it (hipEvents, a few launches each); sets L.shape16/32, L.tuned.
constexpr int kTuned16 = 1, kTuned32 = 2;
// Times every candidate shape of the FP32 16- and/or 32-column SpMM (widths:
// kTuned16 | kTuned32) on this graph and keeps the fastest.
void spmm_autotune(DevLaplacian& L, hipStream_t s, int widths = kTuned16 | kTuned32);
int spmm_shape_count();  // candidates per width
// Average microseconds per launch of the gather probe (the SpMM's 64-byte row
// gathers alone; mode 0 along the column stream, 1 random rows, 2 random rows
// in a 2 MiB footprint), over `reps` launches; see k_spmm.hip.
double gather_probe_us(const DevLaplacian& L, int mode, int reps, hipStream_t s);

// Optional live timing of every SpMM launch (hipEvents on the launching
// stream), used by the benchmark to measure the dominant kernel inside a
// timed solve.  Algorithmic bytes per launch: nnz*(val+4) + 4*(n+1) +
// 2*sizeof(T)*n*k (+ sizeof(T)*n*k when the fused epilogue reads z).
struct SpmmProfile {
    bool enabled = false;
    int64_t launches = 0;
    double bytes = 0.0;       // algorithmic bytes, summed
    double ms = 0.0;          // event time, summed
    double bytes_f32 = 0.0, ms_f32 = 0.0;
    int64_t launches_f32 = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    std::vector<std::pair<double, int>> pending_bytes;  // (bytes, is_f32)
    // events are created once and reused by every later profiled launch, so a
    // profiled region pays two hipEventRecord calls per SpMM and no creation
    std::vector<hipEvent_t> pool;
    size_t pool_next = 0;
};
void spmm_profile_release(SpmmProfile* p);    // destroys the pooled events
void spmm_profile_attach(SpmmProfile* p);     // nullptr detaches
void spmm_profile_collect(SpmmProfile* p);    // sync + fold pending events into the sums

// ----------------------------------------------------------------- dense blocks
// Partial Gram: G = A^T B (FP64 accumulation) for A [n x p], B [n x q].
// Writes the p x q result (row-major, ldg = q) into g; work buffer ws must hold
// gram_workspace_doubles(n, p, q) doubles.
size_t gram_workspace_doubles(int64_t n, int p, int q);
void gram_f32(int64_t n, const float* a, int lda, int p, const float* b, int ldb, int q,
              double* g, double* ws, hipStream_t s);
void gram_f64(int64_t n, const double* a, int lda, int p, const double* b, int ldb, int q,
              double* g, double* ws, hipStream_t s);
// Fused Rayleigh-Ritz update: x = y C, ax = ly C (C m x m), and the sums
// norms2[j] = ||ax_j - theta_j x_j||^2, norms2[k + j] = ||x_j||^2 for j < k <= 8.
// ws: rr_update_workspace_doubles(n, k).
void rr_update_f32(int64_t n, const float* y, const float* ly, int ld, int m, const double* c,
                   const double* theta, int k, float* x, float* ax, double* norms2, double* ws,
                   hipStream_t s);
void rr_update_f64(int64_t n, const double* y, const double* ly, int ld, int m, const double* c,
                   const double* theta, int k, double* x, double* ax, double* norms2, double* ws,
                   hipStream_t s);
size_t rr_update_workspace_doubles(int64_t n, int k);

// ----------------------------------------------------------------- segmented block kernels
// A block-diagonal operator's rows split into segments; a work list of nblk row
// ranges [blk_rows[2b], blk_rows[2b+1]) each inside segment blk_seg[b], grouped
// by segment: segment c owns work items [seg_blk[c], seg_blk[c+1]).  All arrays
// are device arrays.  Gram work items may hold any number of rows; update tiles
// hold at most 64 (FP32) or 32 (FP64) rows.
struct SegWork {
    int nseg = 0;
    int64_t nblk = 0;
    const int64_t* blk_rows = nullptr;
    const int32_t* blk_seg = nullptr;
    const int32_t* seg_blk = nullptr;
};
// Round planner of the batched component solver (k_batch.hip): one workgroup
// reads a round's Ritz values, residual norms and Cholesky status, updates
// every segment's state and writes the next round's per-step, per-segment
// recurrence scalars and the gate ctl[0] (filter steps of the next round, 0 =
// no round).  ctl[1] counts planned rounds, ctl[2] planned SpMM launches,
// ctl[3] the segments still iterating.  Device arrays; theta and thk rows are
// m wide, norms2 rows 2k.
struct BatchPlan {
    int nseg = 0, k = 0, m = 0, cap = 24, steps = 0, max_rounds = 0;
    double tol = 1e-6, b = 2.0;
    double* a = nullptr;      // [nseg] lower end of each segment's damped interval
    double* best = nullptr;   // [nseg] best residual so far (FP32 rounds)
    double* worst = nullptr;  // [nseg] latest residual
    int* stall = nullptr;     // [nseg]
    int* done = nullptr;      // [nseg] 0 iterating, 1 converged (or at the FP32 floor), 2 failed
    int* deg = nullptr;       // [nseg] filter degree of the next round
    double* thk = nullptr;    // [nseg x m] kept Ritz values (FP64 rounds)
    const int* status = nullptr;
    const double* theta = nullptr;
    const double* norms2 = nullptr;
    int32_t* ctl = nullptr;   // [4]
};
// rseg[i] = the segment holding row i, nseg for rows of no segment; the
// segments' row ranges [starts[j], ends[j]) sorted by start, ids[j] their
// segment numbers (device arrays)
// First and last row of every component in the operator's row order: rr[c]
// and rr[nc + c] (device, 2 nc entries; INT32_MAX and -1 for a component
// with no row); perm[row] = input node, comp[node] = component label.
void component_rows(int64_t n, const int32_t* perm, const int32_t* comp, int64_t nc, int32_t* rr,
                    hipStream_t s);
void batch_segment_map(int32_t* rseg, int64_t n, int nseg, const int64_t* starts,
                       const int64_t* ends, const int32_t* ids, hipStream_t s);
// The first filter step of a planned round from the known L X: per row,
// s_l ax + s_x x with its segment's step-1 scalars (coef row 0), into the ring
// buffer a step-1 spmm_seg_ring launch would write; idle when *gate == 0.
void batch_axpby_ring_f32(int64_t n, int w, int ld, const int32_t* rseg, const float* coef,
                          const float* ax, const float* x, float* const ring[3],
                          const int32_t* gate, hipStream_t s);
// dst[r, :k] = src[r, :k] for the rows of every work item (a segment's rows)
void batch_copy_rows(const SegWork& items, const double* src, int lds, double* dst, int ldd,
                     int k, hipStream_t s);
// a = 0.1, best = worst = inf, stall = done = deg = 0, thk = 0, ctl = 0
void batch_plan_init(const BatchPlan& p, hipStream_t s);
// update = 0: plan from the current state only (the first round); 1: read the
// results of the round the gate planned (skipped when that gate was 0); 2: read
// the results of a round launched without a gate
void batch_plan_f32(const BatchPlan& p, int update, float* coef, hipStream_t s);
void batch_plan_f64(const BatchPlan& p, int update, double* coef, hipStream_t s);

// per segment c: G_c = Y_c^T [Y_c | B2_c]  (p x 2p each, g = [nseg x p x 2p]);
// workspace: nblk * 2 p^2 doubles.  The segmented Gram, update and batched
// small-eigenproblem launches take an optional device gate: with *gate == 0
// (a device-planned round with nothing to do) they return at once.
void gram2_seg_f32(const SegWork& blocks, const float* a, int ld, int p, const float* b2, int ld2,
                   double* g, double* ws, hipStream_t s, const int32_t* gate = nullptr);
void gram2_seg_f64(const SegWork& blocks, const double* a, int ld, int p, const double* b2,
                   int ld2, double* g, double* ws, hipStream_t s, const int32_t* gate = nullptr);
// per segment c: x_c = y_c C_c, ax_c = ly_c C_c (C_c = c + c*m*m, theta + c*m),
// norms2 [nseg x 2k]; workspace: nblk * 2k doubles
void rr_update_seg_f32(const SegWork& tiles, const float* y, const float* ly, int ld, int m,
                       const double* c, const double* theta, int k, float* x, float* ax,
                       double* norms2, double* ws, hipStream_t s, const int32_t* gate = nullptr);
void rr_update_seg_f64(const SegWork& tiles, const double* y, const double* ly, int ld, int m,
                       const double* c, const double* theta, int k, double* x, double* ax,
                       double* norms2, double* ws, hipStream_t s, const int32_t* gate = nullptr);
// G = A^T [B1 | B2]  (p x (q1+q2), q1 + q2 <= 64) in one pass over the rows.
void gram2_f32(int64_t n, const float* a, int lda, int p, const float* b, int ldb, int q1,
               const float* b2, int ldb2, int q2, double* g, double* ws, hipStream_t s);
void gram2_f64(int64_t n, const double* a, int lda, int p, const double* b, int ldb, int q1,
               const double* b2, int ldb2, int q2, double* g, double* ws, hipStream_t s);

// out[:, 0:q] = A[:, 0:p] * C (+ B[:, 0:p2] * D), C, D FP64 row-major device matrices.
void block_mul_f32(int64_t n, const float* a, int lda, int p, const double* c, int q,
                   const float* b, int ldb, int p2, const double* d, float* out, int ldo,
                   hipStream_t s);
void block_mul_f64(int64_t n, const double* a, int lda, int p, const double* c, int q,
                   const double* b, int ldb, int p2, const double* d, double* out, int ldo,
                   hipStream_t s);

// r = ax - x * diag(theta) for the first k columns.
void residual_f32(int64_t n, const float* ax, const float* x, int ld, int k, const double* theta,
                  float* r, int ldr, hipStream_t s);
void residual_f64(int64_t n, const double* ax, const double* x, int ld, int k,
                  const double* theta, double* r, int ldr, hipStream_t s);

void fill_normal_f64(double* x, int64_t n, int k, int ld, uint64_t seed, uint32_t stream,
                     hipStream_t s, const int32_t* perm = nullptr);
// FP32 start blocks of the FP32-bulk solvers: the paired FP32 stream
// (normal_pair_f32: one Philox call per row and column pair).
void fill_normal_f32(float* x, int64_t n, int k, int ld, uint64_t seed, uint32_t stream,
                     hipStream_t s, const int32_t* perm = nullptr);
// FP32 block holding normal_at's FP64 draws rounded (the FP64 stream's values).
void fill_normal_rounded_f32(float* x, int64_t n, int k, int ld, uint64_t seed, uint32_t stream,
                             hipStream_t s, const int32_t* perm = nullptr);
// out = alpha a + beta b on columns [0,w) (all three with leading dimension ld)
void axpby_cols_f32(int64_t n, int w, int ld, float alpha, const float* a, float beta,
                    const float* b, float* out, hipStream_t s);
void axpby_cols_f64(int64_t n, int w, int ld, double alpha, const double* a, double beta,
                    const double* b, double* out, hipStream_t s);
void convert_f64_to_f32(const double* a, float* b, int64_t count, hipStream_t s);
void convert_f32_to_f64(const float* a, double* b, int64_t count, hipStream_t s);
// copy columns [0,k) of a [n x lda] block into b [n x ldb]
void copy_cols_f64(const double* a, int lda, double* b, int ldb, int64_t n, int k, hipStream_t s);
void copy_cols_f32(const float* a, int lda, float* b, int ldb, int64_t n, int k, hipStream_t s);
void copy_cols_f64_to_f32(const double* a, int lda, float* b, int ldb, int64_t n, int k,
                          hipStream_t s);
void scale_cols_f32(float* x, int ld, int64_t n, int k, const double* scale, hipStream_t s);
void scale_cols_f64(double* x, int ld, int64_t n, int k, const double* scale, hipStream_t s);
// x[:,c] = (x[:,c] + t[:,c]) / 2
void average_f64(double* x, const double* t, int64_t count, hipStream_t s);
void set_column_f64(double* x, int ld, int64_t n, int col, const double* v, double scale,
                    hipStream_t s);
void set_column_f32(float* x, int ld, int64_t n, int col, const double* v, double scale,
                    hipStream_t s);

// ----------------------------------------------------------------- warm-start network
// out[n x q] = epi(A[n x k] Wa + B[n x k] Wb + bias + C) on the bf16 matrix cores
// (FP32 accumulation); act 0 none, 1 tanh, 2 multiply by (1 - G^2).  (k, q) in
// {(16, 64), (64, 64), (64, 16)}; Wa, Wb bf16 [k x q].
void mpnn_gemm(int64_t n, int k, int q, const float* a, const short* wa, const float* b,
               const short* wb, const float* bias, const float* c, int act, const float* g,
               float* out, hipStream_t s);
void mpnn_dtanh(int64_t count, const float* x, const float* h, float* y, hipStream_t s);
void mpnn_rowfeat(const DevLaplacian& L, double inv_max_sq, float* f, hipStream_t s);
void mpnn_setcol(int64_t n, int j, const float* p, float scale, float* f, hipStream_t s);
void mpnn_deflate(int64_t n, double* p, const double* v, double alpha, float* pf, hipStream_t s);
// output columns of the warm-start network and the coefficients of its loss
// gradient dX = X P + V Q
constexpr int kMpnnOut = 8;
struct MpnnDxCoef {
    double p[kMpnnOut * kMpnnOut];
    double q[2 * kMpnnOut];
};
void mpnn_dx(int64_t n, const float* x, const double* v, const MpnnDxCoef& c, float* dx,
             hipStream_t s);
void mpnn_pack(const float* w, int k, int q, int transpose, short* out, hipStream_t s);
void mpnn_adam(int64_t count, float* p, const double* g, float* m, float* v, float lr, float b1,
               float b2, float eps, int step, hipStream_t s);

// ----------------------------------------------------------------- small dense (one workgroup)
// All small matrices are FP64 row-major, n <= kSmallMax, resident on the device.
constexpr int kSmallMax = 64;
// Largest n of the device dense eigensolver (an n x n FP64 matrix, 128 MiB).
constexpr int kDenseMax = 4096;
// Symmetric eigen-decomposition of a (lower triangle read): w ascending, v columns.
void small_sym_eig(int n, const double* a, int lda, double* w, double* v, hipStream_t s);
// Generalised problem A c = theta B c.  status[0] = 0 ok, 1 if B not SPD.
void small_gen_eig(int n, const double* a, const double* b, double* w, double* c, int* status,
                   hipStream_t s);
// Same for strided A, B, with both sides scaled by diag(B)^-1/2 first; the
// returned vectors apply to the unscaled problem.
void small_gen_eig_scaled(int n, const double* a, int lda, const double* b, int ldb, double* w,
                          double* c, int* status, hipStream_t s);
// count problems of size n at strides sa, sb (doubles); w [count x n],
// c [count x n x n], status [count]
void small_gen_eig_scaled_batched(int count, int n, const double* a, int lda, int64_t sa,
                                  const double* b, int ldb, int64_t sb, double* w, double* c,
                                  int* status, hipStream_t s, const int32_t* gate = nullptr);
// dense EVD of count diagonal blocks (<= 64 rows each) of L: device arrays
// seg_row0 [count], seg_n [count]; evals [count x k]; vectors into rows of evecs
void dense_seg_eig(const DevLaplacian& L, int count, const int64_t* seg_row0, const int32_t* seg_n,
                   int k, double* evals, double* evecs, int ldv, hipStream_t s);
// Given G = X^T X (n x n), writes T = L^-T (upper) so that X T is orthonormal;
// status[0] = 1 when G is not positive definite.
void small_chol_inv(int n, const double* g, double* t, int* status, hipStream_t s);
// scale[c] = 1/sqrt(g[c][c]) (0 when g[c][c] <= 0)
void small_inv_sqrt_diag(int n, const double* g, double* scale, hipStream_t s);
// SVQB orthonormalising transform from the eigenpairs (w, u) of a Gram matrix.
void small_svqb(int n, const double* w, const double* u, double* t, double rel, hipStream_t s);
// Dense FP64 copy of L (n <= kSmallMax), unit diagonal included.
void dense_from_csr(const DevLaplacian& L, double* d, hipStream_t s);
// dst[r * ldd + c] = alpha * src[r * lds + c] for r < rows, c < cols
void small_copy_sub(int rows, int cols, const double* src, int lds, double* dst, int ldd,
                    double alpha, hipStream_t s);
// res[c] = sqrt(g[c][c]) / max(sqrt(gx[c][c]), eps)
void small_residual_norms(int n, const double* gr, int ldr, const double* gx, int ldx,
                          double* res, hipStream_t s);

}  // namespace si
