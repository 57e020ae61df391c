// solver_fast.cpp -- FP32-bulk Chebyshev-filtered block eigensolver.
//
// The reference filters its LOBPCG start block once with a degree-8
// Chebyshev polynomial on [0.1, 2] (lobpcg.rs:113-150, 270-291) and then
// iterates with a block of width n_components + 1.  On clustered k-NN graphs
// the smallest ~#clusters eigenvalues are packed far below the bulk of the
// spectrum, and a block that narrow converges at the rate of the tiny gaps
// inside that cluster.  This solver instead:
//
//   * carries a guarded block of width m (16/32/64, >= the cluster count),
//     so convergence is governed by the gap between the cluster and the bulk;
//   * applies the Chebyshev filter every round, with its damped interval
//     [a, 2] re-centred on the block's largest Ritz value (Zhou-Saad ChFSI);
//   * runs the filter sweeps (the bulk of the bytes) as FP32 SpMMs with the
//     recurrence fused into the SpMM epilogue;
//   * orthonormalises with FP64-accumulated Cholesky-QR (twice) and solves
//     the m x m Rayleigh-Ritz problem in FP64 on the device;
//   * finishes with an FP64 Rayleigh-Ritz and FP64 residual check against the
//     exact Laplacian, adding FP64 filter rounds only if the FP32 floor is
//     above the tolerance.
//
// Convergence is judged on the wanted k = n_components + 1 columns only.
#include <algorithm>
#include <cmath>

#include "si_host.hpp"

namespace si {

namespace {

int choose_block(int64_t n, int k, int requested) {
    int m = requested;
    if (m <= 0) m = n >= 20000 ? 32 : 16;
    const int widths[] = {8, 16, 32, 64};
    int pick = 64;
    for (int w : widths)
        if (w >= m) { pick = w; break; }
    if (pick < k) pick = k <= 8 ? 8 : (k <= 16 ? 16 : (k <= 32 ? 32 : 64));
    while (pick > 8 && pick * 4 > n) pick /= 2;  // keep the block well below n
    return pick;
}

template <typename T>
struct FilterBufs {
    DevBuf<T> b0, b1, b2;
};

// Y = T_d((L - cI)/e) X in precision T, returned in one of the buffers.
template <typename T>
T* cheb_filter(const DevLaplacian& L, FilterBufs<T>& fb, const T* x, int m, double a, double b,
               int degree, int64_t& spmm_calls, hipStream_t s);

template <>
float* cheb_filter<float>(const DevLaplacian& L, FilterBufs<float>& fb, const float* x, int m,
                          double a, double b, int degree, int64_t& spmm_calls, hipStream_t s) {
    const double c = (a + b) / 2.0, e = (b - a) / 2.0;
    const float scale = (float)(2.0 / e), shift = (float)(2.0 * c / e);
    float* yprev = fb.b0.get();
    float* y = fb.b1.get();
    float* ynew = fb.b2.get();
    spmm_f32(L, x, y, m, m, (float)(1.0 / e), (float)(-c / e), 0.f, nullptr, s);
    ++spmm_calls;
    const float* prev = x;
    for (int d = 2; d <= degree; ++d) {
        spmm_f32(L, y, ynew, m, m, scale, -shift, -1.f, prev, s);
        ++spmm_calls;
        float* t = yprev; yprev = y; y = ynew; ynew = t;
        prev = yprev;
    }
    return y;
}

template <>
double* cheb_filter<double>(const DevLaplacian& L, FilterBufs<double>& fb, const double* x, int m,
                            double a, double b, int degree, int64_t& spmm_calls, hipStream_t s) {
    const double c = (a + b) / 2.0, e = (b - a) / 2.0;
    const double scale = 2.0 / e, shift = 2.0 * c / e;
    double* yprev = fb.b0.get();
    double* y = fb.b1.get();
    double* ynew = fb.b2.get();
    spmm_f64(L, x, y, m, m, 1.0 / e, -c / e, 0.0, nullptr, s);
    ++spmm_calls;
    const double* prev = x;
    for (int d = 2; d <= degree; ++d) {
        spmm_f64(L, y, ynew, m, m, scale, -shift, -1.0, prev, s);
        ++spmm_calls;
        double* t = yprev; yprev = y; y = ynew; ynew = t;
        prev = yprev;
    }
    return y;
}

inline void gram_t(int64_t n, const float* a, int lda, int p, const float* b, int ldb, int q,
                   double* g, double* ws, hipStream_t s) {
    gram_f32(n, a, lda, p, b, ldb, q, g, ws, s);
}
inline void gram_t(int64_t n, const double* a, int lda, int p, const double* b, int ldb, int q,
                   double* g, double* ws, hipStream_t s) {
    gram_f64(n, a, lda, p, b, ldb, q, g, ws, s);
}
inline void block_mul_t(int64_t n, const float* a, int lda, int p, const double* c, int q,
                        float* out, int ldo, hipStream_t s) {
    block_mul_f32(n, a, lda, p, c, q, nullptr, 0, 0, nullptr, out, ldo, s);
}
inline void block_mul_t(int64_t n, const double* a, int lda, int p, const double* c, int q,
                        double* out, int ldo, hipStream_t s) {
    block_mul_f64(n, a, lda, p, c, q, nullptr, 0, 0, nullptr, out, ldo, s);
}
inline void scale_cols_t(float* x, int ld, int64_t n, int k, const double* sc, hipStream_t s) {
    scale_cols_f32(x, ld, n, k, sc, s);
}
inline void scale_cols_t(double* x, int ld, int64_t n, int k, const double* sc, hipStream_t s) {
    scale_cols_f64(x, ld, n, k, sc, s);
}
inline void spmm_t(const DevLaplacian& L, const float* x, float* y, int m, hipStream_t s) {
    spmm_f32(L, x, y, m, m, 1.f, 0.f, 0.f, nullptr, s);
}
inline void spmm_t(const DevLaplacian& L, const double* x, double* y, int m, hipStream_t s) {
    spmm_f64(L, x, y, m, m, 1.0, 0.0, 0.0, nullptr, s);
}
inline void residual_t(int64_t n, const float* ax, const float* x, int ld, int k,
                       const double* th, float* r, int ldr, hipStream_t s) {
    residual_f32(n, ax, x, ld, k, th, r, ldr, s);
}
inline void residual_t(int64_t n, const double* ax, const double* x, int ld, int k,
                       const double* th, double* r, int ldr, hipStream_t s) {
    residual_f64(n, ax, x, ld, k, th, r, ldr, s);
}

struct Small {
    DevBuf<double> g, g2, t, w, v, sc, res;
    DevBuf<int> status;
    explicit Small(int m) {
        g.alloc((size_t)m * m); g2.alloc((size_t)m * m); t.alloc((size_t)m * m);
        w.alloc(m); v.alloc((size_t)m * m); sc.alloc(m); res.alloc(m);
        status.alloc(1);
    }
};

// Normalise columns, then Cholesky-QR twice (FP64 Gram).  Returns false if a
// Cholesky factorisation fails.
template <typename T>
bool orthonormalise(int64_t n, T* y, int m, Small& sm, double* ws, hipStream_t s) {
    gram_t(n, y, m, m, y, m, m, sm.g, ws, s);
    small_inv_sqrt_diag(m, sm.g, sm.sc, s);
    scale_cols_t(y, m, n, m, sm.sc, s);
    for (int pass = 0; pass < 2; ++pass) {
        gram_t(n, y, m, m, y, m, m, sm.g, ws, s);
        small_chol_inv(m, sm.g, sm.t, sm.status, s);
        int st = 0;
        to_host(&st, sm.status.get(), 1, s);
        if (st != 0) return false;
        block_mul_t(n, y, m, m, sm.t, m, y, m, s);
    }
    return true;
}

// Rayleigh-Ritz of the block y (not necessarily orthonormal): generalized
// problem (Y^T L Y) c = theta (Y^T Y) c; x <- Y C, ax <- (L Y) C.
template <typename T>
bool rayleigh_ritz(const DevLaplacian& L, int64_t n, T* y, T* ly, int m, T* x, T* ax, Small& sm,
                   double* ws, int64_t& spmm_calls, hipStream_t s) {
    spmm_t(L, y, ly, m, s);
    ++spmm_calls;
    gram_t(n, y, m, m, ly, m, m, sm.g, ws, s);
    gram_t(n, y, m, m, y, m, m, sm.g2, ws, s);
    small_gen_eig(m, sm.g, sm.g2, sm.w, sm.v, sm.status, s);
    int st = 0;
    to_host(&st, sm.status.get(), 1, s);
    if (st != 0) return false;
    block_mul_t(n, y, m, m, sm.v, m, x, m, s);
    block_mul_t(n, ly, m, m, sm.v, m, ax, m, s);
    return true;
}

// residual norms of the first k columns: ||ax - x theta|| / ||x||
template <typename T>
void residual_norms(int64_t n, const T* ax, const T* x, int m, int k, T* r, Small& sm,
                    double* ws, double* res_host, hipStream_t s) {
    residual_t(n, ax, x, m, k, sm.w, r, k, s);
    gram_t(n, r, k, k, r, k, k, sm.g, ws, s);
    gram_t(n, x, m, k, x, m, k, sm.g2, ws, s);
    small_residual_norms(k, sm.g, k, sm.g2, k, sm.res, s);
    to_host(res_host, sm.res.get(), k, s);
}

}  // namespace

bool fast_solve(const DevLaplacian& L, int n_components, uint64_t seed, const FastOptions& opt,
                double* evals_host, double* evecs_dev, SolveStats& st, hipStream_t s) {
    const int64_t n = L.n;
    const int k = n_components + 1;
    st = SolveStats{};
    if (n_components <= 0 || k >= n) return false;
    const int m = choose_block(n, k, opt.block_size);
    if (m >= n) return false;
    st.block_size = m;
    const int degree = opt.cheb_degree > 0 ? opt.cheb_degree : 8;
    const size_t nm = (size_t)n * m;
    DevBuf<double> x64(nm), ws(gram_workspace_doubles(n, m, m)), g(1);
    Small sm(m);

    // start block: warm start (first k columns) or N(0,1) with the trivial hint
    fill_normal_f64(x64, n, m, m, seed, kStreamLobpcgInit, s, L.perm);
    if (opt.warm_start) {
        copy_cols_f64(opt.warm_start, k, x64, m, n, k, s);
    } else {
        gram_f64(n, L.sqrt_deg, 1, 1, L.sqrt_deg, 1, 1, g, ws, s);
        double sn2 = 0.0;
        to_host(&sn2, g.get(), 1, s);
        if (sn2 > 0.0) set_column_f64(x64, m, n, 0, L.sqrt_deg, 1.0 / std::sqrt(sn2), s);
    }

    // ---------------- FP32 rounds
    DevBuf<float> x32(nm), ax32(nm), r32((size_t)n * k);
    FilterBufs<float> fb32;
    fb32.b0.alloc(nm); fb32.b1.alloc(nm); fb32.b2.alloc(nm);
    convert_f64_to_f32(x64, x32, (int64_t)nm, s);
    std::vector<double> theta(m), res(k);
    double a = 0.1;               // first round: the reference's ChFSI interval
    const double b = 2.0;          // spectral radius bound of the normalised Laplacian
    double best = INFINITY;
    int stall = 0;
    bool fp32_ok = false;
    int round = 0;
    // Cap the filter's amplification of the bottom of the spectrum relative to
    // [a, b] at ~5e3, so FP32 noise in the guard columns (1e-7) cannot swamp
    // them after filtering.
    auto degree_for = [&](double lo) {
        const double c = (lo + b) / 2.0, e = (b - lo) / 2.0;
        const double x0 = c / e;  // |(0 - c) / e|
        const int dmax = (int)std::floor(9.2 / std::acosh(std::max(x0, 1.0 + 1e-12)));
        return std::max(2, std::min(degree, dmax));
    };
    for (; round < opt.max_outer; ++round) {
        float* y = cheb_filter<float>(L, fb32, x32, m, a, b, degree_for(a), st.spmm_calls, s);
        if (!orthonormalise<float>(n, y, m, sm, ws, s)) {
            // fall back to a normalised (not orthonormalised) block
            gram_t(n, y, m, m, y, m, m, sm.g, ws, s);
            small_inv_sqrt_diag(m, sm.g, sm.sc, s);
            scale_cols_t(y, m, n, m, sm.sc, s);
        }
        float* ly = (y == fb32.b0.get()) ? fb32.b1.get() : fb32.b0.get();
        if (!rayleigh_ritz<float>(L, n, y, ly, m, x32, ax32, sm, ws, st.spmm_calls, s)) break;
        to_host(theta.data(), sm.w.get(), m, s);
        residual_norms<float>(n, ax32, x32, m, k, r32, sm, ws, res.data(), s);
        ++st.iterations;
        const double worst = *std::max_element(res.begin(), res.end());
        if (worst <= opt.tol) { fp32_ok = true; break; }
        if (worst < 0.7 * best) { best = worst; stall = 0; }
        else if (++stall >= 3 && worst < 1e-4) break;  // FP32 floor reached
        // damp everything at or above the block's top Ritz value
        a = std::min(std::max(theta[m - 1], theta[k - 1] * 1.5 + 1e-12), 0.6);
    }
    (void)fp32_ok;

    // ---------------- FP64 Rayleigh-Ritz, residual and (if needed) FP64 rounds
    DevBuf<double> ax64(nm), r64((size_t)n * k);
    FilterBufs<double> fb64;
    fb64.b0.alloc(nm); fb64.b1.alloc(nm); fb64.b2.alloc(nm);
    convert_f32_to_f64(x32, fb64.b0, (int64_t)nm, s);
    bool ok = rayleigh_ritz<double>(L, n, fb64.b0, fb64.b1, m, x64, ax64, sm, ws, st.spmm_calls, s);
    if (!ok) return false;
    to_host(theta.data(), sm.w.get(), m, s);
    residual_norms<double>(n, ax64, x64, m, k, r64, sm, ws, res.data(), s);
    double worst = *std::max_element(res.begin(), res.end());
    for (int extra = 0; worst > opt.tol && extra < 20; ++extra) {
        a = std::min(std::max(theta[m - 1], theta[k - 1] * 1.5 + 1e-12), 0.6);
        double* y = cheb_filter<double>(L, fb64, x64, m, a, b, degree_for(a), st.spmm_calls, s);
        if (!orthonormalise<double>(n, y, m, sm, ws, s)) break;
        double* ly = (y == fb64.b0.get()) ? fb64.b1.get() : fb64.b0.get();
        if (!rayleigh_ritz<double>(L, n, y, ly, m, x64, ax64, sm, ws, st.spmm_calls, s)) break;
        to_host(theta.data(), sm.w.get(), m, s);
        residual_norms<double>(n, ax64, x64, m, k, r64, sm, ws, res.data(), s);
        ++st.iterations;
        worst = *std::max_element(res.begin(), res.end());
    }
    copy_cols_f64(x64, m, evecs_dev, k, n, k, s);
    SI_HIP(hipStreamSynchronize(s));
    for (int c = 0; c < k; ++c) evals_host[c] = theta[c];
    st.max_residual = worst;
    st.converged = worst <= opt.tol;
    return true;
}

}  // namespace si
