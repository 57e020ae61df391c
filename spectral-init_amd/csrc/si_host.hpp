// si_host.hpp -- host-side runtime of the library: device buffers, the
// context (stream + workspace), timers and the solver entry points shared by
// the C ABI.
#pragma once

#include <chrono>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "si_device.hpp"

namespace si {

// SI_TRACE=1: wall time this thread spent allocating and freeing device
// buffers (DevBuf's hipMalloc / hipFree, StreamBuf's stream-ordered pair).
inline double& devbuf_alloc_ms() {
    thread_local double ms = 0.0;
    return ms;
}

struct AllocTimer {
    std::chrono::steady_clock::time_point t0;
    bool on;
    AllocTimer() : on(trace_on()) {
        if (on) t0 = std::chrono::steady_clock::now();
    }
    ~AllocTimer() {
        if (on)
            devbuf_alloc_ms() += std::chrono::duration<double, std::milli>(
                                     std::chrono::steady_clock::now() - t0).count();
    }
};

// Owning device buffer (hipMalloc'd, freed on destruction).
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        AllocTimer at;
        SI_HIP(hipMalloc(&p, sizeof(T) * count));
        n = count;
    }
    void ensure(size_t count) { if (count > n || p == nullptr) alloc(count); }
    void release() {
        if (p) {
            AllocTimer at;
            hipFree(p);
        }
        p = nullptr;
        n = 0;
    }
    T* get() const { return p; }
    operator T*() const { return p; }
};

// The device's default memory pool keeps what it is given (its release
// threshold raised once per device), so stream-ordered scratch is reused
// across calls instead of being mapped again.
void stream_pool_setup();

// Stream-ordered scratch buffer: hipMallocAsync on stream s, hipFreeAsync on
// the same stream at destruction, so a block is handed out again only to work
// ordered after its last use on s.  For scratch that only work on s touches.
template <typename T>
struct StreamBuf {
    T* p = nullptr;
    size_t n = 0;
    hipStream_t st = nullptr;
    StreamBuf() = default;
    StreamBuf(size_t count, hipStream_t s) { alloc(count, s); }
    StreamBuf(const StreamBuf&) = delete;
    StreamBuf& operator=(const StreamBuf&) = delete;
    ~StreamBuf() { release(); }
    void alloc(size_t count, hipStream_t s) {
        release();
        if (count == 0) count = 1;
        stream_pool_setup();
        AllocTimer at;
        void* q = nullptr;
        SI_HIP(hipMallocAsync(&q, sizeof(T) * count, s));
        p = static_cast<T*>(q);
        n = count;
        st = s;
    }
    void release() {
        if (p) {
            AllocTimer at;
            hipFreeAsync(p, st);
        }
        p = nullptr;
        n = 0;
    }
    T* get() const { return p; }
    operator T*() const { return p; }
};

// Grow-only device workspace, one per host thread (a context is used by one
// host thread).  Solvers take their blocks from named slots, so repeated
// solves allocate nothing after the first one of a given size.  Released
// when the thread exits or by workspace_release().
enum WsSlot : int {
    kWsFastX64, kWsFastAx64, kWsFastR64, kWsFastF64a, kWsFastF64b, kWsFastF64c,
    kWsFastX32, kWsFastAx32, kWsFastR32, kWsFastF32a, kWsFastF32b, kWsFastF32c,
    kWsGram, kWsResAx, kWsResR, kWsRrAx, kWsChainVec, kWsAbiIn, kWsAbiOut,
    kWsSmall0, kWsSmall1, kWsSmall2, kWsSmall3, kWsSmall4, kWsSmall5, kWsSmall6, kWsSmall7,
    kWsPlainS, kWsPlainAS, kWsPlainR, kWsPlainP, kWsPlainAP, kWsPlainPn, kWsPlainAPn,
    kWsPlainX, kWsPlainT, kWsPlainXr, kWsPlainBest, kWsStatus,
    kWsBatchSmall, kWsBatchCoef, kWsBatchRr, kWsBatchPlan, kWsBatchWork, kWsBatchMap,
    kWsDenseA, kWsDenseAux, kWsDensePiv,
    kWsNumSlots
};
void* workspace_get(int slot, size_t bytes);
void workspace_release();
// Per-thread page-locked host buffer of at least `bytes` (grow-only; released
// with the workspace).  Its contents are valid until the next call.
void* pinned_stage(size_t bytes);
// Copy `bytes` of pageable host memory to the device, ordered on stream s.
// Large copies go through a per-thread ring of page-locked chunks (filled by
// host threads while the previous chunk's DMA runs), about twice the rate of a
// pageable copy; src may be reused as soon as the call returns.
void upload_to_device(void* dst, const void* src, size_t bytes, hipStream_t s);
// Copy `bytes` from the device to pageable host memory after the work already
// on stream s, through the same ring of page-locked chunks (each chunk's host
// copy overlaps the next chunks' DMA).  Returns when dst holds the data.
void download_to_host(void* dst, const void* src, size_t bytes, hipStream_t s);
template <typename T>
inline T* ws_get(int slot, size_t count) {
    return static_cast<T*>(workspace_get(slot, sizeof(T) * (count ? count : 1)));
}

// Tuned FP32 SpMM shapes remembered per graph shape (n, stored entries, hub
// rows, widest row, node order): operators of the same shape reuse them.  The
// shapes are performance choices only; every shape computes the same product.
struct TuneKey {
    int64_t n, nnz, n_long, max_row_len;
    int order;
    bool operator<(const TuneKey& o) const {
        if (n != o.n) return n < o.n;
        if (nnz != o.nnz) return nnz < o.nnz;
        if (n_long != o.n_long) return n_long < o.n_long;
        if (max_row_len != o.max_row_len) return max_row_len < o.max_row_len;
        return order < o.order;
    }
};
struct TunedShapes {
    int shape16 = 0, shape32 = 0, tuned = 0;
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    // a second stream (default priority) for small independent work that
    // overlaps the main stream's, joined back through events; made on first use
    hipStream_t aux = nullptr;
    std::string last_error;
    int num_cus = 0;
    SpmmProfile profile;
    std::map<TuneKey, TunedShapes> tune_cache;
};

// The context's auxiliary stream (created on first use).
hipStream_t aux_stream(Context& cx);

// Stream-ordered timer (hipEvents on the context's stream).
struct EventTimer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    explicit EventTimer(hipStream_t st) : s(st) {
        SI_HIP(hipEventCreate(&a));
        SI_HIP(hipEventCreate(&b));
        SI_HIP(hipEventRecord(a, s));
    }
    double stop_ms() {
        SI_HIP(hipEventRecord(b, s));
        SI_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        SI_HIP(hipEventElapsedTime(&ms, a, b));
        return ms;
    }
    ~EventTimer() {
        if (a) hipEventDestroy(a);
        if (b) hipEventDestroy(b);
    }
};

template <typename T>
inline void to_host(T* h, const T* d, size_t count, hipStream_t s) {
    SI_HIP(hipMemcpyAsync(h, d, sizeof(T) * count, hipMemcpyDeviceToHost, s));
    SI_HIP(hipStreamSynchronize(s));
}
template <typename T>
inline void to_device(T* d, const T* h, size_t count, hipStream_t s) {
    SI_HIP(hipMemcpyAsync(d, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
    SI_HIP(hipStreamSynchronize(s));
}

// Result of an eigensolve: host eigenvalues, device eigenvectors [n x k] FP64.
struct SolveStats {
    int level = -1;
    int restarts = 0;
    int block_size = 0;
    int64_t iterations = 0;
    int64_t spmm_calls = 0;
    double max_residual = 0.0;
    bool converged = false;
    bool residual_is_fp64 = false;  // max_residual is the FP64 residual of the returned vectors
    int passes = 0;                 // LOBPCG passes (1 + warm restarts)
    double level_ms[6] = {0, 0, 0, 0, 0, 0};  // wall time per escalation level
};

struct FastOptions {
    int block_size = 0;     // 0 = automatic
    int cheb_degree = 0;    // 0 = automatic: amplification-capped, at most 24
    int max_outer = 60;
    double tol = 1e-6;
    const double* warm_start = nullptr;  // device [n x k]
};

// sqrt(rr) / max(sqrt(vv), 1e-300) from the squared norms rr = ||L v - lambda v||^2
// and vv = ||v||^2 (metrics.rs:57-68); NaN when either is NaN.
double residual_from_sums(double rr, double vv);
// NaN-propagating maximum (max_eigenpair_residual's fold, metrics.rs:80-91).
inline double max_residual_of(const double* r, int k) {
    double q = 0.0;
    for (int c = 0; c < k; ++c) q = (q != q || r[c] != r[c]) ? NAN : (r[c] > q ? r[c] : q);
    return q;
}

// Looks up, once per device and with that device current, the device address
// of the static scratch array of si_max_eigenpair_residual_noalloc.
void residual_noalloc_setup(int device);

// Per-vector FP64 residuals ||L v - lambda v|| / max(||v||, 1e-300) of device
// vectors (metrics.rs:57-92; lobpcg.rs:193-223 batches the same sums).
void fp64_residuals(const DevLaplacian& L, const double* v, int k, const double* evals_host,
                    double* res_host, hipStream_t s);

// Rayleigh-Ritz refinement in FP64 (lobpcg.rs:158-187): v <- v V, evals <- eig(v^T L v).
bool rayleigh_ritz_f64(const DevLaplacian& L, double* v, int k, double* evals_host, hipStream_t s);

// lobpcg_solve (lobpcg.rs:227-465) on the device in FP64 (the plain baseline).
bool lobpcg_solve_plain(const DevLaplacian& L, int n_components, uint64_t seed, bool regularize,
                        double tol, double* evals_host, double* evecs_dev, SolveStats& st,
                        hipStream_t s,
                        const double* warm_start = nullptr);

// The output stage of spectral_init on the device (select_eigenvectors,
// normalize_signs, embed_single_component's placement, scale_and_add_noise):
// b holds the eigenvector block of every segment with its own embedding, each
// segment a contiguous row range [row0, row0 + rows) of it; rowmap / rownode
// map the caller's node ids to rows of b and back (nullptr: the same order).
struct EmbedPlan {
    int64_t n = 0;
    int d = 0;                         // embedding dimension
    const double* b = nullptr;         // device, n x ldb
    int ldb = 0;
    const int32_t* rowmap = nullptr;   // device, node -> row of b
    const int32_t* rownode = nullptr;  // device, row of b -> node
    const int32_t* lab = nullptr;      // device segment label per node (nullptr: one segment)
    int nseg = 1;
    std::vector<int64_t> row0, rows;   // per segment (rows = 0: no block of its own)
    std::vector<int32_t> kind;     // per segment: 1 = own eigenvectors, 0 = meta only
    std::vector<int32_t> cols;     // per segment: d columns of b in selection order
    bool affine = false;           // coordinates sel * (range / max|sel|) + meta
    std::vector<double> range;     // per segment
    std::vector<double> meta;      // per segment, d values
    uint64_t seed = 0;
    float* out = nullptr;          // host, n x d
};
void embed_output_device(const EmbedPlan& p, hipStream_t s);

// Block width of fast_solve for n nodes, k = n_components + 1 wanted pairs and
// a requested width (0: automatic), and the SpMM widths to tune for it.
int fast_block_width(int64_t n, int k, int requested);
inline int fast_tune_widths(int64_t n, int k, int requested) {
    const int m = fast_block_width(n, k, requested);
    return m == 16 ? kTuned16 : (m == 32 ? kTuned32 : 0);
}

// FP32-bulk Chebyshev-filtered block solver; FP64 Rayleigh-Ritz and residual.
bool fast_solve(const DevLaplacian& L, int n_components, uint64_t seed, const FastOptions& opt,
                double* evals_host, double* evecs_dev, SolveStats& st, hipStream_t s);

// Batched FP32-bulk solve of the diagonal blocks of a block-diagonal L (the
// connected components of a graph, stored as contiguous row segments
// [row0[c], row0[c] + rown[c]), each with more than 64 rows).  One launch per
// filter step / Gram / Rayleigh-Ritz update serves every segment; each segment
// has its own filter interval and degree.  evals_host [nseg x k]; evecs_dev
// [L.n x k] receives each segment's vectors in its rows (rows outside every
// segment are left untouched); resid [nseg] the FP64
// residual of each segment's pairs; ok[c] = 0 if segment c did not reach tol.
struct BatchStats {
    int rounds = 0;
    int64_t spmm_calls = 0;
    int graphs_captured = 0;
    int graph_launches = 0;
};
void fast_solve_segments(const DevLaplacian& L, const std::vector<int64_t>& row0,
                         const std::vector<int64_t>& rown, int n_components, uint64_t seed,
                         const FastOptions& opt, bool use_graphs, double* evals_host,
                         double* evecs_dev, std::vector<double>& resid, std::vector<int>& ok,
                         BatchStats& st, hipStream_t s);

// chebyshev_precond (lobpcg.rs:113-150) in FP64 on device blocks.
void chebyshev_filter_f64(const DevLaplacian& L, const double* r, int k, double a, double b,
                          int degree, double* out, hipStream_t s);

// rsvd_solve (rsvd.rs:135-213).
void rsvd_solve_device(const DevLaplacian& L, int n_components, uint64_t seed, int oversampling,
                       int nbiter, double* evals_host, double* evecs_dev, SolveStats& st,
                       hipStream_t s);

// dense_evd (dense.rs:9-46) for n <= kSmallMax (one-workgroup Jacobi).
void dense_evd_small(const DevLaplacian& L, int k, double* evals_host, double* evecs_dev,
                     hipStream_t s);
// dense_evd for n <= kDenseMax (Householder tridiagonalisation, multisection,
// inverse iteration, back-transformation).
void dense_evd_large(const DevLaplacian& L, int k, double* evals_host, double* evecs_dev,
                     hipStream_t s);
// dense_evd for any n <= kDenseMax: the Jacobi path up to kSmallMax nodes,
// the tridiagonal path above.
void dense_evd_device(const DevLaplacian& L, int k, double* evals_host, double* evecs_dev,
                      hipStream_t s);

// A symmetric block operator for lobpcg_core: out = A in, k columns, leading
// dimension ld for both blocks, enqueued on the solver's stream.
using BlockOperator = std::function<void(const double* in, double* out, int k, int ld)>;
// Knyazev's block LOBPCG (the algorithm of linfa_linalg::lobpcg) for the m
// smallest eigenpairs of op; see solver_plain.cpp.
// robust: normalise the active residual columns before their Cholesky-QR and,
// if it still fails, lock the most converged active column and retry, where
// the plain method stops.
bool lobpcg_core(int64_t n, const BlockOperator& op, double* x, int m, double tol, int maxiter,
                 double* lam_h, double* rnorm_h, int64_t& iters, int64_t& op_calls,
                 hipStream_t s, bool robust = false);

// Banded Cholesky factor of M = L + shift I in the device's node order
// (k_band.hip).  The band is bounded by kBandMaxEntries stored doubles and
// kBandMaxFlops for the factorisation; larger bands are not factorised.
constexpr double kBandMaxEntries = 134217728.0;  // 1 GiB of FP64
constexpr double kBandMaxFlops = 4e12;
// When the band exceeds those limits, M is factorised as a dense matrix
// instead (rocSOLVER DPOTRF on the lower triangle, column-major n x n) if
// n <= kSinvDenseMax and n^2 doubles take at most kSinvDenseMemFrac of the free
// device memory; the solves are two rocBLAS DTRSM on the row-major block.
constexpr int64_t kSinvDenseMax = 65536;
constexpr double kSinvDenseMemFrac = 0.6;
struct BandFactor {
    int n = 0;
    int b = 0;           // bandwidth: M[i][j] == 0 for i - j > b (n - 1 when dense)
    bool dense = false;  // B holds the dense Cholesky factor (lower, column-major)
    DevBuf<double> B;
};
int band_width(const DevLaplacian& L, hipStream_t s);
// false when neither a band nor a dense factor is available (too large) or M
// is not positive definite
bool band_factor(const DevLaplacian& L, double shift, BandFactor& F, hipStream_t s);
// testing: factorise densely whatever the band (the dense fallback's seam)
void sinv_force_dense(bool on);
// y = scale * M^-1 x (scale +1 or -1) for k columns
void band_solve(const BandFactor& F, const double* x, int ldx, double* y, int ldy, int k,
                double scale, hipStream_t s);

// lobpcg_sinv_solve (sinv.rs:106-189): LOBPCG on (L + 1e-4 I)^-1; false for
// the reference's None (band too large, factorisation failed, no convergence).
constexpr double kSinvShift = 1e-4;             // SINV_SHIFT (sinv.rs:18)
constexpr double kSinvQualityThreshold = 1e-6;  // SINV_LOBPCG_QUALITY_THRESHOLD (metrics.rs:35)
bool lobpcg_sinv_solve_device(const DevLaplacian& L, int n_components, uint64_t seed,
                              double* evals_host, double* evecs_dev, SolveStats& st,
                              hipStream_t s);

// solve_eigenproblem (solvers/mod.rs:70-224).
struct ChainOptions {
    int solver = 0;              // si_solver_kind at level 1
    double tol = 1e-5;
    int64_t dense_threshold = 2000;
    FastOptions fast;
};
int solve_eigenproblem_device(const DevLaplacian& L, int n_components, uint64_t seed,
                              const ChainOptions& opt, double* evals_host, double* evecs_dev,
                              SolveStats& st, hipStream_t s);

}  // namespace si
