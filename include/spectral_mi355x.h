/*
 * spectral_mi355x.h -- C ABI of the MI355X spectral-initialisation library.
 *
 * Drop-in boundary for spectral-init's eigensolver path.  Each entry point
 * names the reference item it replaces (paths relative to the reference
 * crate).  All functions return an si_status; no C++ exception or panic
 * crosses this boundary.  si_last_error() returns the message of the last
 * failing call on a context (the Display text of the reference's errors).
 *
 * Pointer conventions: arguments documented "host" are CPU memory; arguments
 * documented "device" are HIP device memory on the context's device.  Dense
 * blocks are row-major [n x k].  CSR input: int64 row pointers (the
 * reference's usize), int32 column indices (u32), float weights (f32).
 * One context per host thread; a context owns one HIP stream and its
 * workspaces.
 */
#ifndef SPECTRAL_MI355X_H
#define SPECTRAL_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* SpectralError (src/lib.rs:167-182) plus device / argument failures. */
typedef enum {
    SI_OK = 0,
    SI_ERR_CONVERGENCE = 1,      /* SpectralError::ConvergenceFailure */
    SI_ERR_INVALID_GRAPH = 2,    /* SpectralError::InvalidGraph(String) */
    SI_ERR_TOO_FEW_NODES = 3,    /* SpectralError::TooFewNodes { n, dims } */
    SI_ERR_HIP = 4,              /* HIP runtime failure */
    SI_ERR_INTERNAL = 5,         /* an internal invariant failed (the reference panics) */
    SI_ERR_INVALID_ARGUMENT = 6, /* bad pointer / size / option */
    SI_ERR_UNSUPPORTED = 7       /* outside the supported envelope (e.g. nnz >= 2^31) */
} si_status;

/* ComputeMode (src/config.rs:3-14). */
typedef enum {
    SI_COMPUTE_PYTHON_COMPAT = 0,
    SI_COMPUTE_RUST_NATIVE = 1
} si_compute_mode;

/* Which eigensolver runs at escalation level 1. */
typedef enum {
    SI_SOLVER_FAST = 0,          /* FP32-bulk Chebyshev-filtered block solver, FP64 RR + residual */
    SI_SOLVER_PLAIN_LOBPCG = 1   /* FP64 LOBPCG with the reference's constants (the baseline) */
} si_solver_kind;

typedef enum { SI_F32 = 0, SI_F64 = 1 } si_precision;

/* SpectralInitConfig (src/config.rs:19-25) plus the device knobs.  Set
 * struct_size = sizeof(si_config); si_config_default() gives the reference's
 * behaviour (PythonCompat, tol 1e-5, dense threshold 2000). */
typedef struct {
    uint32_t struct_size;
    int32_t compute_mode;      /* si_compute_mode */
    int32_t solver;            /* si_solver_kind */
    int32_t block_size;        /* fast solver block width: 0 = automatic, else 8/16/32/64 */
    int32_t cheb_degree;       /* fast solver: max Chebyshev degree per round, 0 = automatic (<= 24);
                                  the plain LOBPCG keeps the reference prefilter degree 8 (lobpcg.rs:72) */
    int32_t max_outer;         /* fast solver: cap on filter/Rayleigh-Ritz rounds */
    double tol;                /* residual tolerance (reference LOBPCG_ACCEPT_TOL 1e-5) */
    int64_t dense_threshold;   /* level-0 dense EVD below this n (reference 2000, solvers/mod.rs:32) */
    const double* warm_start;  /* optional device block [n x (n_components+1)] to start from */
    int32_t batch_components;  /* disconnected graphs: 1 = all components in one batched device
                                  solve (default), 0 = one solve per component as the reference
                                  loops (src/multi_component.rs:261-312) */
} si_config;

/* Per-call report (timing and quality, the reference's [timing:level_N]
 * and metrics.rs residual). */
typedef struct {
    uint32_t struct_size;
    int32_t level;             /* escalation level that produced the result (solvers/mod.rs:70) */
    int32_t restarts;          /* LOBPCG restart_count (lobpcg.rs:300-306) */
    int32_t block_size;        /* block width the solver used */
    int64_t iterations;        /* LOBPCG iterations or filter rounds */
    int64_t spmm_calls;        /* applications of L (any width) */
    double max_residual;       /* max_j ||L v_j - lambda_j v_j|| / ||v_j|| in FP64 (metrics.rs:73) */
    double eigenvalues[8];     /* first min(k, 8) eigenvalues */
    double solve_ms;           /* device time of the eigensolve */
    double laplacian_ms;       /* device time of the Laplacian build */
    /* fields below are written only when struct_size covers them */
    int32_t passes;            /* LOBPCG passes run: 1 + warm restarts of either kind (from
                                  partial eigenvectors or after a failed residual check,
                                  lobpcg.rs:315-450); 0 for the other solvers */
    int32_t reserved;
    double level_ms[6];        /* wall time of each escalation level tried, L0..L5, 0 for
                                  levels not reached: the reference's [timing:level_N] lines
                                  (solvers/mod.rs:76-223; also printed with SI_TRACE=1) */
} si_report;

typedef struct si_context si_context;
typedef struct si_laplacian si_laplacian;

const char* si_version(void);
void si_config_default(si_config* cfg);
si_status si_context_create(int device, si_context** out);
void si_context_destroy(si_context* ctx);
const char* si_last_error(const si_context* ctx);

/* compute_degrees + build_normalized_laplacian (src/laplacian.rs:29-88),
 * producing a device-resident operator.  input_on_device = 0: indptr /
 * indices / data are host arrays; 1: device arrays.  Graphs with n >= 4096
 * are stored in a locality node order (see SI_BUILD_LOCALITY_ORDER); every
 * entry point below takes and returns vectors in the caller's node order. */
si_status si_laplacian_build(si_context* ctx, int64_t n, const int64_t* indptr,
                             const int32_t* indices, const float* data, int input_on_device,
                             int32_t compute_mode, si_laplacian** out);
/* Same, with explicit build flags: SI_BUILD_LOCALITY_ORDER stores the graph in
 * a locality order (a connected symmetric graph of >= 65536 nodes with >= 8
 * stored entries per node: rows grouped by k-means clusters of a short
 * random-walk embedding; otherwise reverse Cuthill-McKee), SI_BUILD_RCM_ORDER
 * in reverse Cuthill-McKee order always, 0 in the input order. */
#define SI_BUILD_LOCALITY_ORDER 1u
#define SI_BUILD_RCM_ORDER 2u
si_status si_laplacian_build_ex(si_context* ctx, int64_t n, const int64_t* indptr,
                                const int32_t* indices, const float* data, int input_on_device,
                                int32_t compute_mode, uint32_t flags, si_laplacian** out);
void si_laplacian_destroy(si_laplacian* lap);
/* n, stored entries of W, number of hub rows (> 256 entries), and whether the
 * PythonCompat degrees reproduce the f32 column-sum order exactly. */
si_status si_laplacian_info(const si_laplacian* lap, int64_t* n, int64_t* nnz, int64_t* n_long,
                            int32_t* degree_exact);
/* Copies degrees / sqrt_deg (host, n each) and L in the reference's CSR form
 * (host: l_indptr[n+1], l_indices[nnz_L], l_data[nnz_L], unit diagonal
 * included, columns ascending); any output may be NULL.  nnz_L is returned. */
si_status si_laplacian_export(si_context* ctx, const si_laplacian* lap, double* degrees,
                              double* sqrt_deg, int64_t* l_indptr, int32_t* l_indices,
                              double* l_data, int64_t* nnz_l);

/* LinearOperator::apply / CsrOperator::apply (src/operator.rs:6-12, 176-228):
 * y = L x for device blocks of width 1 <= k <= 4096 and leading dimension
 * ld >= k (widths 16/32/64 take the vector path when ld and the pointers are
 * 16-byte aligned; other widths are split into supported column chunks).  SI_F32 uses FP32 values (FP32 accumulation); SI_F64 the
 * exact FP64 Laplacian. */
si_status si_spmm(si_context* ctx, const si_laplacian* lap, int32_t precision, const void* x,
                  void* y, int32_t k, int32_t ld);

/* LinearOperator::apply (src/operator.rs:6-12) on host arrays: y = L x for
 * row-major host blocks [n x k] (FP64, the caller's node order), computed on
 * the device with the exact FP64 Laplacian.  The entry a Rust
 * `impl LinearOperator` forwards to. */
si_status si_operator_apply_host(si_context* ctx, const si_laplacian* lap, const double* x,
                                 double* y, int32_t k);

/* spmv_csr (src/operator.rs:16-57), the fixed raw-slice "SIMD replacement
 * point": y[i] = sum_k data[k] * x[indices[k]] for a general CSR matrix given
 * as host slices (usize row pointers and column indices, f64 values), computed
 * on the device with the scalar loop's rounding (each product rounded, then
 * added in storage order), so y equals the reference's bit for bit.
 * indptr has n_rows + 1 entries, indices and data at least indptr[n_rows];
 * x has x_len entries, y n_rows.  The reference's precondition panics
 * (indptr length, column bounds, data length) return SI_ERR_INVALID_ARGUMENT. */
si_status si_spmv_csr(si_context* ctx, int64_t n_rows, const uint64_t* indptr,
                      const uint64_t* indices, const double* data, int64_t data_len,
                      const double* x, int64_t x_len, double* y);

/* chebyshev_precond (src/solvers/lobpcg.rs:113-150): out = T_d((L - cI)/e) r
 * with c = (a+b)/2, e = (b-a)/2; FP64 device blocks [n x k]. */
si_status si_chebyshev_filter(si_context* ctx, const si_laplacian* lap, const double* r,
                              int32_t k, double a, double b, int32_t degree, double* out);

/* max_eigenpair_residual (src/metrics.rs:73-92) in FP64 for device
 * eigenvectors [n x k] and host eigenvalues[k]. */
si_status si_max_eigenpair_residual(si_context* ctx, const si_laplacian* lap,
                                    const double* evals, const double* evecs, int32_t k,
                                    double* out);

/* The same residual as si_max_eigenpair_residual, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  It has no counterpart in the reference.
 * The call makes no hipMalloc / hipFree: its partial sums live in a static
 * device array of 1 MiB + 1 KiB per GPU, held until the process ends and
 * shared by every call on that GPU, so calls that overlap on one GPU race.
 * evals (host, k), evecs (device, n x k row-major, the caller's node order),
 * out (host); k <= 64, larger k is rejected with SI_ERR_INVALID_ARGUMENT. */
si_status si_max_eigenpair_residual_noalloc(si_context* ctx, const si_laplacian* lap,
                                            const double* evals, const double* evecs, int32_t k,
                                            double* out);

/* lobpcg_solve (src/solvers/lobpcg.rs:227-465) on the device, FP64, with the
 * reference constants; tol as LOBPCG_ACCEPT_TOL.  evals (host, k =
 * n_components+1), evecs (device, n x k).  Returns SI_ERR_CONVERGENCE for
 * the reference's None. */
si_status si_lobpcg_solve(si_context* ctx, const si_laplacian* lap, int32_t n_components,
                          uint64_t seed, int32_t regularize, double tol, double* evals,
                          double* evecs, si_report* rep);

/* lobpcg_sinv_solve (src/solvers/sinv.rs:106-189): LOBPCG on (L + 1e-4 I)^-1
 * with the reference's start block, tolerance (1e-8) and iteration cap
 * (min(5n, 300)); eigenvalues lambda = 1/mu - 1e-4, ascending.  L + 1e-4 I is
 * factorised by a banded Cholesky in the stored node order (reverse
 * Cuthill-McKee for n >= 4096); SI_ERR_CONVERGENCE for the reference's None,
 * which includes a band above 2^27 stored entries or 4e12 factorisation
 * flops.  evals host, evecs device (n x (n_components+1)). */
si_status si_lobpcg_sinv_solve(si_context* ctx, const si_laplacian* lap, int32_t n_components,
                               uint64_t seed, double* evals, double* evecs, si_report* rep);

/* solve_eigenproblem (src/solvers/mod.rs:70-224): escalation L0 dense EVD
 * (n < dense_threshold), L1 (cfg->solver), L2 shift-and-invert LOBPCG, L3
 * regularised LOBPCG, L4 rSVD, L5 dense EVD (n <= 4096), each gated on the
 * FP64 residual.  evals host, evecs device. */
si_status si_solve_eigenproblem(si_context* ctx, const si_laplacian* lap, int32_t n_components,
                                uint64_t seed, const si_config* cfg, double* evals,
                                double* evecs, si_report* rep);

/* rsvd_solve (src/solvers/rsvd.rs:135-213): randomized SVD of 2I - L with
 * oversampling (< 0: the reference's rsvd_k_sub) and nbiter power
 * iterations; evals host, evecs device. */
si_status si_rsvd_solve(si_context* ctx, const si_laplacian* lap, int32_t n_components,
                        uint64_t seed, int32_t oversampling, int32_t nbiter, double* evals,
                        double* evecs, si_report* rep);

/* dense_evd (src/solvers/dense.rs:9-46): the k smallest eigenpairs of L, FP64,
 * for n <= 4096 and k <= 64: one-workgroup Jacobi up to n = 64, Householder
 * tridiagonalisation + Sturm multisection + inverse iteration above.  evals
 * host (ascending), evecs device (n x k).  SI_ERR_UNSUPPORTED above n = 4096. */
si_status si_dense_evd(si_context* ctx, const si_laplacian* lap, int32_t k, double* evals,
                       double* evecs);

/* spectral_init (src/lib.rs:198-258): host CSR graph in, host
 * out[n x n_components] f32 coordinates.  raw_data (host, n x n_features f32)
 * is needed only for disconnected graphs with more than 2*n_components
 * components (src/multi_component.rs:31-48). */
si_status si_spectral_init(si_context* ctx, int64_t n, const int64_t* indptr,
                           const int32_t* indices, const float* data, int32_t n_components,
                           uint64_t seed, const float* raw_data, int64_t n_features,
                           const si_config* cfg, float* out, si_report* rep);

#ifdef __cplusplus
}
#endif
#endif
