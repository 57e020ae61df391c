/*
 * spectral_mi355x_testing.h -- kernel-level seams for parity tests.
 *
 * The reference exposes testing-only internals behind its `testing` cargo
 * feature (src/lib.rs:27-162; e.g. spmv_csr as the fixed-signature SIMD
 * replacement point, src/operator.rs:16-32).  These entry points are the
 * device counterparts: they drive one kernel family at a time so the tests can
 * compare it with the CPU oracle.  Not part of the stable ABI.
 */
#ifndef SPECTRAL_MI355X_TESTING_H
#define SPECTRAL_MI355X_TESTING_H

#include "spectral_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

/* G = A^T B with FP64 accumulation; a, b device [n x lda] / [n x ldb]; g host [p x q]. */
si_status si_test_gram(si_context* ctx, int32_t precision, int64_t n, const void* a, int32_t lda,
                       int32_t p, const void* b, int32_t ldb, int32_t q, double* g_host);
/* out = A C; a, out device; c host [p x q] FP64. */
si_status si_test_block_mul(si_context* ctx, int32_t precision, int64_t n, const void* a,
                            int32_t lda, int32_t p, const double* c_host, int32_t q, void* out,
                            int32_t ldo);
/* one-workgroup Jacobi eigen-decomposition (host in / out, n <= 64). */
si_status si_test_small_sym_eig(si_context* ctx, int32_t n, const double* a_host, double* w_host,
                                double* v_host);
/* one-workgroup generalised eigenproblem A c = theta B c (host in / out). */
si_status si_test_small_gen_eig(si_context* ctx, int32_t n, const double* a_host,
                                const double* b_host, double* w_host, double* c_host);
/* Philox4x32-10 / Box-Muller normals into a device block [n x k]. */
si_status si_test_fill_normal(si_context* ctx, double* x, int64_t n, int32_t k, uint64_t seed,
                              uint32_t stream);
/* the FP32 start-block stream of the FP32-bulk solvers into a device float
 * block [n x k]: row i, column pair p holds pair i * ceil(k/2) + p (cosine in
 * column 2p, sine in column 2p + 1). */
si_status si_test_fill_normal_f32(si_context* ctx, float* x, int64_t n, int32_t k, uint64_t seed,
                                  uint32_t stream);
/* host bytes to the device and back through the library's staged copies
 * (page-locked chunk ring for >= 8 MiB, the runtime's pageable copy below):
 * out = in, byte for byte, via a device buffer of `bytes`. */
si_status si_test_staged_roundtrip(si_context* ctx, const void* in, void* out, int64_t bytes);
/* the FP32-bulk Chebyshev-filtered block solver alone (escalation level 1). */
si_status si_test_fast_solve(si_context* ctx, const si_laplacian* lap, int32_t n_components,
                             uint64_t seed, const si_config* cfg, double* evals, double* evecs,
                             si_report* rep);

/* y = (L + shift I)^-1 x for k device columns (caller's node order) through
 * the banded Cholesky of the shift-and-invert level (src/solvers/sinv.rs:125,
 * 139-153); *bandwidth receives the band used (-1 when unavailable). */
si_status si_test_band_solve(si_context* ctx, const si_laplacian* lap, double shift,
                             const double* x, double* y, int32_t k, int32_t* bandwidth);

/* Shift-and-invert level: enable != 0 factorises M = L + shift I as a dense
 * matrix (rocSOLVER Cholesky) whatever its band, so the dense fallback used for
 * graphs with no narrow band can be compared with the banded factor; 0 restores
 * the automatic choice (band first, dense when the band is too large). */
si_status si_test_sinv_dense(int32_t enable);

/* Batched component solves (si_spectral_init with batch_components): enable
 * >= 0 switches hipGraph capture of the FP32 rounds on (1) or off (0), -1
 * leaves it; captured / launches (may be NULL) receive the graph counts of
 * the last batched solve. */
si_status si_test_batch_graphs(int32_t enable, int64_t* captured, int64_t* launches);

/* One call of the batched solver's device round planner (k_batch.hip) on
 * given round results, all arrays on the host.  mode 0: the FP32 rule, 1: the
 * FP64 rule; update 0 plans from the state only, 1 reads the round the gate
 * ctl[0] planned (skipped when it is 0), 2 always reads it.  In / out: the
 * segment state a, best, worst [nseg] (doubles), stall, done [nseg], thk
 * [nseg x m] (mode 1 only; may be NULL in mode 0), ctl [4] (gate, planned
 * rounds, planned SpMM launches, iterating segments).  In: status [nseg],
 * theta [nseg x m], norms2 [nseg x 2k].  Out: deg [nseg] and the recurrence
 * scalars coef [steps x (nseg + 1) x 3], widened to double in mode 0. */
si_status si_test_batch_plan(si_context* ctx, int32_t mode, int32_t update, int32_t nseg,
                             int32_t k, int32_t m, int32_t cap, int32_t steps,
                             int32_t max_rounds, double tol, double* a, double* best,
                             double* worst, int32_t* stall, int32_t* done, double* thk,
                             const int32_t* status, const double* theta, const double* norms2,
                             int32_t* ctl, int32_t* deg, double* coef);

/* FP32 SpMM shape for 16 / 32 columns (0 .. 3, see k_spmm.hip): autotune != 0
 * times every candidate on this graph and keeps the fastest, otherwise the
 * given shapes are set; out16 / out32 (may be NULL) receive the shapes used. */
si_status si_test_spmm_shapes(si_context* ctx, si_laplacian* lap, int32_t shape16,
                              int32_t shape32, int32_t autotune, int32_t* out16, int32_t* out32);

/* Live timing of every SpMM launch on the context's stream (hipEvents around
 * each launch) between begin and end; bytes are the algorithmic bytes of
 * each launch (nnz*(value+index) + row pointers + x read + y write). */
typedef struct {
    int64_t launches;
    double bytes;
    double ms;
    int64_t launches_f32;
    double bytes_f32;
    double ms_f32;
} si_spmm_profile;
si_status si_profile_spmm_begin(si_context* ctx);
si_status si_profile_spmm_end(si_context* ctx, si_spmm_profile* out);

/* The SpMM's gather stream alone, as a measure of the rate the memory system
 * gives it: one 64-byte FP32 row (16 columns) read per stored entry of the
 * Laplacian, no values, row pointers or output rows.  mode 0 reads the rows
 * the column array names (the SpMM's own gathers), mode 1 random rows over
 * the same n rows, mode 2 random rows among the first 32768 (2 MiB, resident
 * in every L2).  *us receives the average microseconds per launch over reps
 * launches. */
si_status si_profile_gather_probe(si_context* ctx, const si_laplacian* lap, int32_t mode,
                                  int32_t reps, double* us);

/* The locality order of si_laplacian_build (reverse Cuthill-McKee; perm[new] =
 * old) of a host CSR pattern, on `threads` host threads (0: as many as the
 * hardware allows, at most 16).  Host-only: needs no device. */
si_status si_test_rcm_order(int64_t n, const int64_t* indptr, const int32_t* indices,
                            int32_t threads, int32_t* perm);

/* Where si_spectral_init forms its output in later calls of this process
 * (selection, signs, component placement, scaling, noise): 1 = on the device
 * (the default), 0 = on the host.  Both give the same bits. */
si_status si_test_output_stage(int32_t device);

/* The device counterparts of the two host steps above, for a symmetric host
 * CSR pattern (uploaded here): component labels (numbered by smallest node, as
 * find_components) with their count, and the locality order perm[new] = old. */
si_status si_test_components_device(si_context* ctx, int64_t n, const int64_t* indptr,
                                    const int32_t* indices, int32_t* labels, int64_t* n_comp);
si_status si_test_rcm_order_device(si_context* ctx, int64_t n, const int64_t* indptr,
                                   const int32_t* indices, int32_t* perm);

/* The operator as stored on the device: the CSR of W in the storage node order
 * (rowptr[n+1], col[nnz], the FP32 sweep values a32[nnz] = +W_ij / sqrt(d_i d_j)),
 * perm[n] (perm[new] = old; the identity when no locality order was applied)
 * and which order it is (0 input, 1 reverse Cuthill-McKee, 2 clustering).  Any
 * output may be NULL. */
si_status si_test_laplacian_storage(si_context* ctx, const si_laplacian* lap, int32_t* rowptr,
                                    int32_t* col, float* a32, int32_t* perm, int32_t* order);

/* The arithmetic of the exact k-NN step (si_knn_exact, si_fuzzy_graph_build)
 * in later calls of this process: mode 0 = automatic (points with k <= 32
 * take the matrix-core filter, bf16x3 or FP32, and the certified FP64
 * re-rank), 1 = the direct FP64 kernel only, 2 = the filter with every query
 * re-run by the direct kernel, 3 = the filter on the FP32 matrix cores only;
 * -1 leaves the mode unchanged.  last_direct (may be null)
 * receives the number of queries the last filtered call re-ran directly.
 * Results are identical in every mode. */
si_status si_test_knn_path(int32_t mode, int64_t* last_direct);

#ifdef __cplusplus
}
#endif
#endif
