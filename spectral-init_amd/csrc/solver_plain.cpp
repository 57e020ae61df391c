// solver_plain.cpp -- the plain GPU LOBPCG: a direct device form of the
// reference's lobpcg_solve (src/solvers/lobpcg.rs:227-465) in FP64.
//
// Block width k = n_components + 1, N(0,1) start with the trivial-vector
// hint, one ChFSI filter (a = 0.1, b = 2, degree 8) for n >= 1000, up to three
// warm restarts, Rayleigh-Ritz refinement and the per-vector residual gate.
// The LOBPCG iteration is Knyazev's block method with soft locking (the
// algorithm of linfa_linalg::lobpcg, which the reference calls at
// lobpcg.rs:327-358): residuals, projection against X, Cholesky-QR of the
// active residuals and of P, Rayleigh-Ritz on span[X, R, P].  One kernel
// launch per operation and a host-side convergence check every iteration --
// this is the baseline the FP32-bulk solver is measured against.
#include <cmath>
#include <cstdio>
#include <functional>

#include "si_host.hpp"

namespace si {

namespace {

struct PlainWork {
    int64_t n = 0;
    int m = 0;
    // n-sized blocks come from the thread's workspace pool
    double *S, *AS, *R, *P, *AP, *Pn, *APn, *ws;
    DevBuf<double> g1, g2, t1, w, c, csub, sel, eye, lam, colsc;
    DevBuf<int> status;
    PlainWork(int64_t n_, int m_) : n(n_), m(m_) {
        const int s3 = 3 * m;
        S = ws_get<double>(kWsPlainS, (size_t)n * s3);
        AS = ws_get<double>(kWsPlainAS, (size_t)n * s3);
        R = ws_get<double>(kWsPlainR, (size_t)n * m);
        P = ws_get<double>(kWsPlainP, (size_t)n * m);
        AP = ws_get<double>(kWsPlainAP, (size_t)n * m);
        Pn = ws_get<double>(kWsPlainPn, (size_t)n * m);
        APn = ws_get<double>(kWsPlainAPn, (size_t)n * m);
        ws = ws_get<double>(kWsGram, gram_workspace_doubles(n, s3, s3));
        g1.alloc((size_t)s3 * s3);
        g2.alloc((size_t)s3 * s3);
        t1.alloc((size_t)s3 * s3);
        w.alloc(s3);
        c.alloc((size_t)s3 * s3);
        csub.alloc((size_t)s3 * m);
        sel.alloc((size_t)m * m);
        eye.alloc((size_t)m * m);
        lam.alloc(m);
        colsc.alloc(m);
        status.alloc(1);
    }
};

int read_status(const DevBuf<int>& st, hipStream_t s) {
    int h = 0;
    to_host(&h, st.get(), 1, s);
    return h;
}

void upload_eye(DevBuf<double>& eye, int m, hipStream_t s) {
    std::vector<double> h((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) h[i * m + i] = 1.0;
    to_device(eye.get(), h.data(), h.size(), s);
}

}  // namespace

// One LOBPCG pass for the m smallest eigenpairs of the symmetric operator op
// (op(in, out, k, ld): out = A in for k columns, both with leading dimension
// ld).  x: device [n x m] (ld m), replaced by the final Ritz vectors;
// lam_h/rnorm_h receive the Ritz values (ascending) and residual norms.
bool lobpcg_core(int64_t n, const BlockOperator& op, double* x, int m, double tol, int maxiter,
                 double* lam_h, double* rnorm_h, int64_t& iters, int64_t& op_calls,
                 hipStream_t s, bool robust) {
    const int s3 = 3 * m;
    // the Rayleigh-Ritz problems on span[X, R] (and P) are solved in one
    // workgroup, at most kSmallMax wide: blocks wider than half of that are
    // not run here (the escalation chain moves on), and P is left out of an
    // iteration whose span[X, R, P] would not fit
    if (m > kSmallMax / 2) return false;
    PlainWork W(n, m);
    upload_eye(W.eye, m, s);
    double* X = W.S;
    double* AX = W.AS;
    auto apply = [&](const double* in, double* out, int k) {
        op(in, out, k, s3);
        ++op_calls;
    };
    copy_cols_f64(x, m, X, s3, n, m, s);
    // orthonormalise X (Cholesky-QR)
    gram_f64(n, X, s3, m, X, s3, m, W.g1, W.ws, s);
    small_chol_inv(m, W.g1, W.t1, W.status, s);
    if (read_status(W.status, s) != 0) return false;
    block_mul_f64(n, X, s3, m, W.t1, m, nullptr, 0, 0, nullptr, X, s3, s);
    apply(X, AX, m);
    gram_f64(n, X, s3, m, AX, s3, m, W.g1, W.ws, s);
    small_sym_eig(m, W.g1, m, W.w, W.c, s);
    block_mul_f64(n, X, s3, m, W.c, m, nullptr, 0, 0, nullptr, X, s3, s);
    block_mul_f64(n, AX, s3, m, W.c, m, nullptr, 0, 0, nullptr, AX, s3, s);
    SI_HIP(hipMemcpyAsync(W.lam.get(), W.w.get(), sizeof(double) * m, hipMemcpyDeviceToDevice, s));
    to_host(lam_h, W.w.get(), m, s);

    bool has_p = false, converged = false;
    int64_t it = 0;
    std::vector<double> hg((size_t)s3 * s3);
    std::vector<int> act(m);
    for (;;) {
        residual_f64(n, AX, X, s3, m, W.lam, W.R, m, s);
        gram_f64(n, W.R, m, m, W.R, m, m, W.g1, W.ws, s);
        to_host(hg.data(), W.g1.get(), (size_t)m * m, s);
        int na = 0;
        for (int c = 0; c < m; ++c) {
            rnorm_h[c] = std::sqrt(std::fmax(hg[c * m + c], 0.0));
            act[c] = rnorm_h[c] > tol;
            na += act[c];
        }
        if (trace_on() && (it % 25 == 0 || it >= maxiter)) {
            std::fprintf(stderr, "[si] lobpcg it %lld active %d rnorm", (long long)it, na);
            for (int c = 0; c < m; ++c) std::fprintf(stderr, " %.3e", rnorm_h[c]);
            std::fprintf(stderr, "\n");
        }
        if (na == 0) { converged = true; break; }
        if (it >= maxiter) break;
        double* Ra = X + m;
        double* ARa = AX + m;
        bool r_ok = false;
        for (;;) {
            // selection matrix of the active columns
            std::vector<double> hs((size_t)m * na, 0.0);
            for (int c = 0, q = 0; c < m; ++c)
                if (act[c]) hs[c * na + q++] = 1.0;
            to_device(W.sel.get(), hs.data(), hs.size(), s);
            block_mul_f64(n, W.R, m, m, W.sel, na, nullptr, 0, 0, nullptr, Ra, s3, s);
            // R_a <- R_a - X (X^T R_a)
            gram_f64(n, X, s3, m, Ra, s3, na, W.g1, W.ws, s);
            small_copy_sub(m, na, W.g1, na, W.g2, na, -1.0, s);
            std::vector<double> he((size_t)na * na, 0.0);
            for (int i = 0; i < na; ++i) he[i * na + i] = 1.0;
            to_device(W.t1.get(), he.data(), he.size(), s);
            block_mul_f64(n, Ra, s3, na, W.t1, na, X, s3, m, W.g2, Ra, s3, s);
            // Cholesky-QR of the active residuals.  In the robust mode their
            // columns are first scaled to unit norm (the same span and Q in
            // exact arithmetic, a unit-diagonal Gram matrix), and when the
            // residuals are numerically dependent -- as for the two members of
            // an exactly degenerate pair -- the most converged active column is
            // locked and the factorisation retried.
            gram_f64(n, Ra, s3, na, Ra, s3, na, W.g1, W.ws, s);
            if (robust) {
                std::vector<double> hd((size_t)na * na), hsc(na);
                to_host(hd.data(), W.g1.get(), hd.size(), s);
                for (int c = 0; c < na; ++c) {
                    const double d = hd[(size_t)c * na + c];
                    hsc[c] = d > 1e-300 ? 1.0 / std::sqrt(d) : 1.0;
                }
                to_device(W.colsc.get(), hsc.data(), na, s);
                scale_cols_f64(Ra, s3, n, na, W.colsc, s);
                gram_f64(n, Ra, s3, na, Ra, s3, na, W.g1, W.ws, s);
            }
            small_chol_inv(na, W.g1, W.t1, W.status, s);
            if (read_status(W.status, s) == 0) { r_ok = true; break; }
            if (trace_on())
                std::fprintf(stderr, "[si] lobpcg it %lld: residual Cholesky failed (%d active)\n",
                             (long long)it, na);
            if (!robust || na <= 1) break;
            int drop = -1;
            for (int c = 0; c < m; ++c)
                if (act[c] && (drop < 0 || rnorm_h[c] < rnorm_h[drop])) drop = c;
            act[drop] = 0;
            --na;
        }
        if (!r_ok) break;
        block_mul_f64(n, Ra, s3, na, W.t1, na, nullptr, 0, 0, nullptr, Ra, s3, s);
        apply(Ra, ARa, na);
        bool use_p = false;
        if (has_p && m + 2 * na <= kSmallMax) {
            double* Pa = X + m + na;
            double* APa = AX + m + na;
            block_mul_f64(n, W.P, m, m, W.sel, na, nullptr, 0, 0, nullptr, Pa, s3, s);
            block_mul_f64(n, W.AP, m, m, W.sel, na, nullptr, 0, 0, nullptr, APa, s3, s);
            gram_f64(n, Pa, s3, na, Pa, s3, na, W.g1, W.ws, s);
            small_chol_inv(na, W.g1, W.t1, W.status, s);
            if (read_status(W.status, s) == 0) {
                block_mul_f64(n, Pa, s3, na, W.t1, na, nullptr, 0, 0, nullptr, Pa, s3, s);
                block_mul_f64(n, APa, s3, na, W.t1, na, nullptr, 0, 0, nullptr, APa, s3, s);
                use_p = true;
            }
        }
        int sdim = 0;
        bool ok = false;
        for (int attempt = 0; attempt < 2 && !ok; ++attempt) {
            if (attempt == 1) {
                if (!use_p) break;
                use_p = false;
            }
            sdim = m + na + (use_p ? na : 0);
            gram_f64(n, X, s3, sdim, AX, s3, sdim, W.g1, W.ws, s);
            gram_f64(n, X, s3, sdim, X, s3, sdim, W.g2, W.ws, s);
            small_gen_eig(sdim, W.g1, W.g2, W.w, W.c, W.status, s);
            ok = read_status(W.status, s) == 0;
        }
        if (!ok) {
            if (trace_on()) std::fprintf(stderr, "[si] lobpcg it %lld: Rayleigh-Ritz failed\n", (long long)it);
            break;
        }
        small_copy_sub(sdim, m, W.c, sdim, W.csub, m, 1.0, s);
        const double* csub_rp = W.csub.get() + (size_t)m * m;
        block_mul_f64(n, X + m, s3, sdim - m, csub_rp, m, nullptr, 0, 0, nullptr, W.Pn, m, s);
        block_mul_f64(n, AX + m, s3, sdim - m, csub_rp, m, nullptr, 0, 0, nullptr, W.APn, m, s);
        block_mul_f64(n, X, s3, m, W.csub, m, W.Pn, m, m, W.eye, X, s3, s);
        block_mul_f64(n, AX, s3, m, W.csub, m, W.APn, m, m, W.eye, AX, s3, s);
        std::swap(W.P, W.Pn);
        std::swap(W.AP, W.APn);
        SI_HIP(hipMemcpyAsync(W.lam.get(), W.w.get(), sizeof(double) * m, hipMemcpyDeviceToDevice, s));
        to_host(lam_h, W.w.get(), m, s);
        has_p = true;
        ++it;
    }
    copy_cols_f64(X, s3, x, m, n, m, s);
    SI_HIP(hipStreamSynchronize(s));
    iters += it;
    return converged;
}

namespace {

bool lobpcg_core_plain(const DevLaplacian& L, double shift, double* x, int m, double tol,
                       int maxiter, double* lam_h, double* rnorm_h, int64_t& iters,
                       int64_t& spmm_calls, hipStream_t s) {
    const BlockOperator op = [&](const double* in, double* out, int k, int ld) {
        spmm_f64(L, in, out, k, ld, 1.0, shift, 0.0, nullptr, s);
    };
    return lobpcg_core(L.n, op, x, m, tol, maxiter, lam_h, rnorm_h, iters, spmm_calls, s);
}

}  // namespace

bool lobpcg_solve_plain(const DevLaplacian& L, int n_components, uint64_t seed, bool regularize,
                        double tol, double* evals_host, double* evecs_dev, SolveStats& st,
                        hipStream_t s, const double* warm_start) {
    const int64_t n = L.n;
    const int k = n_components + 1;
    st = SolveStats{};
    st.block_size = k;
    if (n_components <= 0 || k >= n) return false;  // lobpcg.rs:239-241
    const size_t nk = (size_t)n * k;
    DevBuf<double> x(nk), t(nk), g(1), ws(gram_workspace_doubles(n, k, k)), gk((size_t)k * k),
        sc(k);
    if (warm_start) {
        // a caller's start block takes the place of the random one
        SI_HIP(hipMemcpyAsync(x.get(), warm_start, sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
    } else {
        // N(0,1) start (lobpcg.rs:244-246) and the trivial-vector hint (:249-252)
        fill_normal_f64(x, n, k, k, seed, kStreamLobpcgInit, s, L.perm);
        gram_f64(n, L.sqrt_deg, 1, 1, L.sqrt_deg, 1, 1, g, ws, s);
        double sn2 = 0.0;
        to_host(&sn2, g.get(), 1, s);
        const double sn = std::sqrt(sn2);
        if (sn > 0.0) set_column_f64(x, k, n, 0, L.sqrt_deg, 1.0 / sn, s);
    }
    if (n >= 1000) {  // ChFSI (lobpcg.rs:270-291)
        chebyshev_filter_f64(L, x, k, 0.1, 2.0, 8, t, s);
        st.spmm_calls += 8;
        average_f64(x, t, (int64_t)nk, s);
        gram_f64(n, x, k, k, x, k, k, gk, ws, s);
        std::vector<double> h((size_t)k * k), hs(k);
        to_host(h.data(), gk.get(), h.size(), s);
        for (int c = 0; c < k; ++c) {
            const double nr = std::sqrt(std::fmax(h[c * k + c], 0.0));
            hs[c] = nr > 1e-300 ? 1.0 / nr : 1.0;
        }
        to_device(sc.get(), hs.data(), k, s);
        scale_cols_f64(x, k, n, k, sc, s);
    }
    const double shift = regularize ? 1e-5 : 0.0;  // REGULARIZATION_EPS (lobpcg.rs:59)
    std::vector<double> lam(k), rn(k), res(k), best_l(k);
    DevBuf<double> xr(nk), best;
    bool have_last = false;
    int restart_count = 0;
    for (int restart = 0; restart <= 3; ++restart) {
        const int64_t cap = restart == 0 ? 300 : 1000;  // lobpcg.rs:89-97
        const int maxiter = (int)((5 * n) < cap ? 5 * n : cap);
        ++st.passes;
        const bool conv = lobpcg_core_plain(L, shift, x, k, tol, maxiter, lam.data(), rn.data(),
                                            st.iterations, st.spmm_calls, s);
        bool accept = conv;
        if (!conv) {
            accept = true;
            for (int c = 0; c < k; ++c)
                if (!(rn[c] < tol)) accept = false;
        }
        if (!accept) {
            if (restart < 3) continue;  // warm restart from the partial eigvecs
            break;
        }
        SI_HIP(hipMemcpyAsync(xr.get(), x.get(), sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
        if (!rayleigh_ritz_f64(L, xr, k, lam.data(), s)) {
            if (restart < 3) continue;
            break;
        }
        st.spmm_calls += 1;
        fp64_residuals(L, xr, k, lam.data(), res.data(), s);
        st.spmm_calls += 1;
        bool all_ok = true;
        for (int c = 0; c < k; ++c)
            if (!(res[c] < tol)) all_ok = false;
        if (all_ok) {
            SI_HIP(hipMemcpyAsync(evecs_dev, xr.get(), sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
            SI_HIP(hipStreamSynchronize(s));
            for (int c = 0; c < k; ++c) evals_host[c] = lam[c];
            st.converged = true;
            st.restarts = restart_count;
            return true;
        }
        ++restart_count;
        best.ensure(nk);
        SI_HIP(hipMemcpyAsync(best.get(), xr.get(), sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
        best_l = lam;
        have_last = true;
        SI_HIP(hipMemcpyAsync(x.get(), xr.get(), sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
    }
    st.restarts = restart_count;
    if (!have_last) return false;
    SI_HIP(hipMemcpyAsync(evecs_dev, best.get(), sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
    SI_HIP(hipStreamSynchronize(s));
    for (int c = 0; c < k; ++c) evals_host[c] = best_l[c];
    return true;
}

}  // namespace si
