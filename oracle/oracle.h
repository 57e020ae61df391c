/*
 * oracle.h -- FP64 CPU restatement of spectral-init's eigensolver hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against (and the CPU baseline timed beside it).  Only the
 * tests, the smoke check and the benchmark's cpu_baseline leg may load it; the
 * product library never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Conventions: dense blocks are row-major [n x k]; CSR row pointers are int64,
 * column indices int32 (the reference uses usize/u32; values fit).
 *
 * Parity pinning (what checks each function independently of this code):
 *   or_compute_degrees / or_build_laplacian  -> reference fixtures comp_a / comp_b
 *   or_spmm                                  -> dense numpy product
 *   or_chebyshev_filter                      -> numpy chebval on diagonal operators
 *   or_sym_eig / or_dense_evd                -> numpy.linalg.eigh
 *   or_rayleigh_ritz                         -> rotated-basis closed form
 *   or_lobpcg_solve / or_solve_eigenproblem  -> scipy eigsh fixtures comp_d, ring
 *                                               closed form 1-cos(2*pi*j/n), Rust
 *                                               eigenvector blocks (subspace angle)
 *   or_rsvd_solve                            -> ring-6 closed form, numpy eigh
 *   or_find_components / selection / signs  -> fixtures comp_c / comp_e / comp_f
 *   or_mpnn_*                                -> central finite differences
 *   or_knn_exact .. or_fuzzy_union           -> fixtures step1_knn_exact ..
 *                                               step5a_pruned_exact
 */
#ifndef SI_ORACLE_H
#define SI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- threading (OpenMP); 1 = faithful single-threaded reference ---- */
void or_set_threads(int n);
int or_get_threads(void);

/* ---- counter-based normal RNG (Philox4x32-10 + Box-Muller), shared by the
 *      oracle and the device path so both start from the same blocks ---- */
void or_philox4x32_10(const uint32_t ctr[4], uint32_t key0, uint32_t key1, uint32_t out[4]);
double or_normal(uint64_t seed, uint32_t stream, uint64_t index);
void or_fill_normal(uint64_t seed, uint32_t stream, size_t count, double* out);
/* pairs p = 0 .. npairs-1 of the FP32 stream: out[2p] (cosine), out[2p+1] (sine) */
void or_normal_pairs_f32(uint64_t seed, uint32_t stream, size_t npairs, float* out);

/* ---- laplacian.rs ---- */
/* mode 0 = PythonCompat (f32 column-sum scatter, laplacian.rs:8-16,37-40),
 * mode 1 = RustNative (f64 row sums, laplacian.rs:41-46). */
void or_compute_degrees(int64_t n, const int64_t* indptr, const int32_t* indices,
                        const float* data, int mode, double* deg, double* sqrt_deg);
/* nnz of L = nnz(W) - (#explicit diagonal entries of W) + n (laplacian.rs:60-88). */
int64_t or_laplacian_nnz(int64_t n, const int64_t* indptr, const int32_t* indices);
void or_build_laplacian(int64_t n, const int64_t* indptr, const int32_t* indices,
                        const float* data, const double* inv_sqrt_deg,
                        int64_t* l_indptr, int32_t* l_indices, double* l_data);

/* ---- operator.rs: y = A x for row-major blocks (CsrOperator::apply, :176-228) ---- */
void or_spmm(int64_t n, const int64_t* indptr, const int32_t* indices, const double* data,
             const double* x, int k, double* y);

/* ---- metrics.rs:57-92 ---- */
double or_eigenpair_residual(int64_t n, const int64_t* indptr, const int32_t* indices,
                             const double* data, const double* v, double lambda);
double or_max_eigenpair_residual(int64_t n, const int64_t* indptr, const int32_t* indices,
                                 const double* data, const double* vecs /* n x k row-major */,
                                 const double* lambdas, int k);

/* ---- small dense FP64 kernels ---- */
int or_cholesky(int n, double* a /* n x n row-major, lower on exit */);
void or_sym_eig(int n, const double* a, double* w /* ascending */, double* v /* n x n, cols */);
int or_gen_sym_eig(int n, const double* a, const double* b, double* w, double* c);
/* dense_evd (solvers/dense.rs:9-46): k smallest eigenpairs of the densified CSR. */
int or_dense_evd(int64_t n, const int64_t* indptr, const int32_t* indices, const double* data,
                 int k, double* evals, double* evecs /* n x k */);

/* ---- lobpcg.rs ---- */
/* chebyshev_precond (lobpcg.rs:113-150): returns T_d((L-cI)/e) r. */
void or_chebyshev_filter(int64_t n, const int64_t* indptr, const int32_t* indices,
                         const double* data, const double* r, int k, double a, double b,
                         int degree, double* out);
/* rayleigh_ritz_refine (lobpcg.rs:158-187); x is replaced by X*V, returns 0 on success. */
int or_rayleigh_ritz(int64_t n, const int64_t* indptr, const int32_t* indices,
                     const double* data, double* x, int k, double* evals);
/* per_vector_residuals (lobpcg.rs:193-223). */
void or_per_vector_residuals(int64_t n, const int64_t* indptr, const int32_t* indices,
                             const double* data, const double* evals, const double* x, int k,
                             double* res);

typedef struct {
    int64_t iterations;      /* LOBPCG iterations summed over all passes */
    int64_t spmm_calls;      /* operator applications */
    int restarts;            /* restart_count as returned by lobpcg_solve */
    int passes;              /* LOBPCG passes run (initial + warm restarts of any kind) */
    int converged;           /* 1 if every per-vector residual < tol */
} or_lobpcg_stats;

/* Core block LOBPCG iteration (the arithmetic of linfa_linalg::lobpcg called at
 * lobpcg.rs:327-358): smallest m eigenpairs, soft locking, X/R/P Rayleigh-Ritz.
 * x: in = initial block, out = Ritz vectors; returns 1 if all rnorm < tol. */
int or_lobpcg_core(int64_t n, const int64_t* indptr, const int32_t* indices,
                   const double* data, double shift, double* x, int m, double tol,
                   int maxiter, double* evals, double* rnorm, int64_t* iters,
                   int64_t* spmm_calls);

/* lobpcg_solve (lobpcg.rs:227-465) with the reference constants (ChFSI a=0.1,
 * b=2, d=8 when n >= 1000; <= 3 warm restarts; maxiter min(5n,300)/min(5n,1000)).
 * tol is LOBPCG_ACCEPT_TOL (1e-5 in the reference, metrics.rs:46).
 * Returns 1 when a result is produced (Some), 0 for None. */
int or_lobpcg_solve(int64_t n, const int64_t* indptr, const int32_t* indices,
                    const double* data, int n_components, uint64_t seed, int regularize,
                    const double* sqrt_deg, double tol, const double* x0_or_null,
                    double* evals, double* evecs, or_lobpcg_stats* st);

/* rsvd_solve (rsvd.rs:135-213); oversampling < 0 selects rsvd_k_sub (rsvd.rs:84-87). */
void or_rsvd_solve(int64_t n, const int64_t* indptr, const int32_t* indices,
                   const double* data, int n_components, uint64_t seed, int oversampling,
                   int nbiter, double* evals, double* evecs);

/* solve_eigenproblem escalation (solvers/mod.rs:70-224). Level 2 (sparse
 * Cholesky shift-invert) is not restated; the chain goes L0, L1, L3, L4, L5.
 * Returns the level that produced the result. */
int or_solve_eigenproblem(int64_t n, const int64_t* indptr, const int32_t* indices,
                          const double* data, int n_components, uint64_t seed,
                          const double* sqrt_deg, int64_t dense_threshold, double lobpcg_tol,
                          double* evals, double* evecs, or_lobpcg_stats* st);

/* ---- components.rs / selection.rs / scaling.rs ---- */
int64_t or_find_components(int64_t n, const int64_t* indptr, const int32_t* indices,
                           int64_t* labels);
void or_select_eigenvectors(const double* evals, const double* evecs, int64_t n, int k,
                            int n_components, double* out);
void or_normalize_signs(double* coords, int64_t n, int d);
int or_scale_coords(const double* coords, int64_t n, int d, double max_coord, float* out);

/* ---- spectral_init (lib.rs:198-258) up to the noise draw; noise uses the
 *      counter RNG above (stream 7).  Returns 0 on success, else SI error code. */
int or_spectral_init(int64_t n, const int64_t* indptr, const int32_t* indices,
                     const float* data, int n_components, uint64_t seed,
                     const float* raw_data, int64_t n_features, int mode,
                     int64_t dense_threshold, double lobpcg_tol, int add_noise,
                     float* out);

/* ---- fuzzy k-NN graph (the input W; oracle_graph.c) ----
 * Fixture pipeline tests/generate_fixtures.py:70-182 (umap-learn 0.5.11). */
void or_knn_exact(int64_t n, int d, const double* x, int k, int32_t* idx, float* dist);
void or_smooth_knn_dist(int64_t n, int k, const float* dist, float* sigmas, float* rhos);
void or_membership(int64_t n, int k, const int32_t* idx, const float* dist, const float* sigmas,
                   const float* rhos, float* vals);
int64_t or_fuzzy_union(int64_t n, int k, const int32_t* idx, const float* vals, int n_epochs,
                       int64_t* indptr, int32_t* indices, float* data);

/* ---- learned warm start (no reference counterpart; oracle_mpnn.c) ---- */
int64_t or_mpnn_param_count(void);
void or_mpnn_features(int64_t n, const int64_t* ip, const int32_t* ix, const double* a,
                      const double* deg, uint64_t seed, double* f);
void or_mpnn_forward(int64_t n, const int64_t* ip, const int32_t* ix, const double* a,
                     const double* f, const double* params, double* h, double* av, double* x);
double or_mpnn_loss(int64_t n, const double* x, const double* v, double* dx);
void or_mpnn_backward(int64_t n, const int64_t* ip, const int32_t* ix, const double* a,
                      const double* f, const double* params, const double* h,
                      const double* av, const double* dx, double* grad);

#ifdef __cplusplus
}
#endif
#endif
