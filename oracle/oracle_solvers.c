/*
 * oracle_solvers.c -- Chebyshev filter, Rayleigh-Ritz refinement, block
 * LOBPCG, the lobpcg_solve driver, randomized SVD and the escalation chain.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline, never the
 * product).  The LOBPCG iteration restates Knyazev's block LOBPCG (SIAM J. Sci.
 * Comput. 23(2), 2001) with soft locking -- the published algorithm behind
 * linfa_linalg::lobpcg, which the reference calls at
 * src/solvers/lobpcg.rs:327-358 and which is not vendored there.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define STREAM_LOBPCG_INIT 1u
#define STREAM_RSVD_OMEGA 2u

typedef struct {
    int64_t n;
    const int64_t* indptr;
    const int32_t* indices;
    const double* data;
    double shift; /* y = L x + shift x (Level 3 regularisation, lobpcg.rs:333) */
    int64_t calls;
} op_t;

static void op_apply(op_t* op, const double* x, int k, double* y) {
    or_spmm(op->n, op->indptr, op->indices, op->data, x, k, y);
    if (op->shift != 0.0) {
        size_t cnt = (size_t)op->n * k;
        for (size_t i = 0; i < cnt; ++i) y[i] += op->shift * x[i];
    }
    op->calls++;
}

/* G = A^T B for row-major A [n x p], B [n x q]. */
static void gram(int64_t n, const double* a, int p, const double* b, int q, double* g) {
    memset(g, 0, sizeof(double) * (size_t)p * q);
    for (int64_t i = 0; i < n; ++i) {
        const double* ai = a + (size_t)i * p;
        const double* bi = b + (size_t)i * q;
        for (int r = 0; r < p; ++r) {
            double ar = ai[r];
            for (int c = 0; c < q; ++c) g[r * q + c] += ar * bi[c];
        }
    }
}

/* out [n x q] = A [n x p] * C [p x q] */
static void block_mul(int64_t n, const double* a, int p, const double* c, int q, double* out) {
    for (int64_t i = 0; i < n; ++i) {
        const double* ai = a + (size_t)i * p;
        double* oi = out + (size_t)i * q;
        for (int col = 0; col < q; ++col) {
            double s = 0.0;
            for (int r = 0; r < p; ++r) s += ai[r] * c[r * q + col];
            oi[col] = s;
        }
    }
}

/* X <- X * L^-T where L L^T = X^T X (Cholesky-QR).  Also applies the same
 * transform to companion block ax when given.  Returns -1 on failure. */
static int cholqr(int64_t n, double* x, int k, double* ax) {
    double* g = (double*)malloc(sizeof(double) * (size_t)k * k);
    gram(n, x, k, x, k, g);
    if (or_cholesky(k, g) != 0) { free(g); return -1; }
    /* inverse of L^T (upper), column-wise: Uinv = (L^T)^-1 */
    double* uinv = (double*)calloc((size_t)k * k, sizeof(double));
    for (int col = 0; col < k; ++col) {
        for (int i = k - 1; i >= 0; --i) {
            double s = (i == col) ? 1.0 : 0.0;
            for (int p = i + 1; p < k; ++p) s -= g[p * k + i] * uinv[p * k + col];
            uinv[i * k + col] = s / g[i * k + i];
        }
    }
    double* tmp = (double*)malloc(sizeof(double) * (size_t)n * k);
    block_mul(n, x, k, uinv, k, tmp);
    memcpy(x, tmp, sizeof(double) * (size_t)n * k);
    if (ax) {
        block_mul(n, ax, k, uinv, k, tmp);
        memcpy(ax, tmp, sizeof(double) * (size_t)n * k);
    }
    free(tmp); free(uinv); free(g);
    return 0;
}

/* chebyshev_precond (src/solvers/lobpcg.rs:113-150), same operation order. */
void or_chebyshev_filter(int64_t n, const int64_t* indptr, const int32_t* indices,
                         const double* data, const double* r, int k, double a, double b,
                         int degree, double* out) {
    size_t cnt = (size_t)n * k;
    double c = (a + b) / 2.0, e = (b - a) / 2.0;
    double scale = 2.0 / e, shift = scale * c;
    double* yprev = (double*)malloc(sizeof(double) * cnt);
    double* y = (double*)malloc(sizeof(double) * cnt);
    double* ly = (double*)malloc(sizeof(double) * cnt);
    double* ynew = (double*)malloc(sizeof(double) * cnt);
    memcpy(yprev, r, sizeof(double) * cnt);
    or_spmm(n, indptr, indices, data, r, k, ly);
    for (size_t i = 0; i < cnt; ++i) y[i] = (ly[i] - c * r[i]) / e;
    for (int d = 2; d <= degree; ++d) {
        or_spmm(n, indptr, indices, data, y, k, ly);
        for (size_t i = 0; i < cnt; ++i) {
            double t = scale * ly[i];
            t -= shift * y[i];
            t -= yprev[i];
            ynew[i] = t;
        }
        double* t = yprev; yprev = y; y = ynew; ynew = t;
    }
    memcpy(out, y, sizeof(double) * cnt);
    free(yprev); free(y); free(ly); free(ynew);
}

/* rayleigh_ritz_refine (src/solvers/lobpcg.rs:158-187). */
static int rr_refine(op_t* op, double* x, int k, double* evals) {
    int64_t n = op->n;
    double* ax = (double*)malloc(sizeof(double) * (size_t)n * k);
    double* g = (double*)malloc(sizeof(double) * (size_t)k * k);
    double* v = (double*)malloc(sizeof(double) * (size_t)k * k);
    op_t plain = *op;
    plain.shift = 0.0; /* refinement uses the true Laplacian */
    op_apply(&plain, x, k, ax);
    op->calls++;
    gram(n, x, k, ax, k, g);
    int ok = 1;
    for (int i = 0; i < k * k; ++i) if (!isfinite(g[i])) ok = 0;
    if (ok) {
        or_sym_eig(k, g, evals, v);
        block_mul(n, x, k, v, k, ax);
        memcpy(x, ax, sizeof(double) * (size_t)n * k);
    }
    free(ax); free(g); free(v);
    return ok ? 0 : -1;
}

int or_rayleigh_ritz(int64_t n, const int64_t* indptr, const int32_t* indices,
                     const double* data, double* x, int k, double* evals) {
    op_t op = {n, indptr, indices, data, 0.0, 0};
    return rr_refine(&op, x, k, evals);
}

/* per_vector_residuals (src/solvers/lobpcg.rs:193-223). */
void or_per_vector_residuals(int64_t n, const int64_t* indptr, const int32_t* indices,
                             const double* data, const double* evals, const double* x, int k,
                             double* res) {
    double* ax = (double*)malloc(sizeof(double) * (size_t)n * k);
    or_spmm(n, indptr, indices, data, x, k, ax);
    for (int c = 0; c < k; ++c) {
        double xn = 0.0, dn = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            double xv = x[(size_t)i * k + c];
            double d = ax[(size_t)i * k + c] - evals[c] * xv;
            xn += xv * xv;
            dn += d * d;
        }
        xn = sqrt(xn);
        if (xn < 2.220446049250313e-16) xn = 2.220446049250313e-16;
        res[c] = sqrt(dn) / xn;
    }
    free(ax);
}

/* ------------------------------------------------------------------------ */
/* Block LOBPCG (smallest eigenpairs) with soft locking.                       */
static int lobpcg_core_op(op_t* op, double* x, int m, double tol, int maxiter, double* lam,
                          double* rnorm, int64_t* iters) {
    int64_t n = op->n;
    size_t nm = (size_t)n * m;
    double* ax = (double*)malloc(sizeof(double) * nm);
    double* r = (double*)malloc(sizeof(double) * nm);
    double* p = (double*)calloc(nm, sizeof(double));
    double* ap = (double*)calloc(nm, sizeof(double));
    int smax = 3 * m;
    double* s = (double*)malloc(sizeof(double) * (size_t)n * smax);
    double* as = (double*)malloc(sizeof(double) * (size_t)n * smax);
    double* ga = (double*)malloc(sizeof(double) * (size_t)smax * smax);
    double* gb = (double*)malloc(sizeof(double) * (size_t)smax * smax);
    double* th = (double*)malloc(sizeof(double) * (size_t)smax);
    double* cc = (double*)malloc(sizeof(double) * (size_t)smax * smax);
    double* csub = (double*)malloc(sizeof(double) * (size_t)smax * m);
    double* xr = (double*)malloc(sizeof(double) * (size_t)m * m);
    int* act = (int*)malloc(sizeof(int) * (size_t)m);
    int has_p = 0, converged = 0;
    int64_t it = 0;

    /* initial orthonormalisation + Rayleigh-Ritz */
    if (cholqr(n, x, m, NULL) != 0) goto done;
    op_apply(op, x, m, ax);
    gram(n, x, m, ax, m, ga);
    or_sym_eig(m, ga, lam, xr);
    block_mul(n, x, m, xr, m, r); memcpy(x, r, sizeof(double) * nm);
    block_mul(n, ax, m, xr, m, r); memcpy(ax, r, sizeof(double) * nm);

    for (;;) {
        /* residuals and active set */
        for (size_t i = 0; i < nm; ++i) r[i] = ax[i] - x[i] * lam[i % m];
        int na = 0;
        for (int c = 0; c < m; ++c) {
            double q = 0.0;
            for (int64_t i = 0; i < n; ++i) q += r[(size_t)i * m + c] * r[(size_t)i * m + c];
            rnorm[c] = sqrt(q);
            act[c] = rnorm[c] > tol;
            na += act[c];
        }
        if (na == 0) { converged = 1; break; }
        if (it >= maxiter) break;
        /* gather active residual columns into s[:, m : m+na] layout later */
        double* ra = (double*)malloc(sizeof(double) * (size_t)n * na);
        double* ara = (double*)malloc(sizeof(double) * (size_t)n * na);
        for (int64_t i = 0; i < n; ++i) {
            int q = 0;
            for (int c = 0; c < m; ++c) if (act[c]) ra[(size_t)i * na + q++] = r[(size_t)i * m + c];
        }
        /* project out X, orthonormalise */
        double* xtr = (double*)malloc(sizeof(double) * (size_t)m * na);
        gram(n, x, m, ra, na, xtr);
        for (int64_t i = 0; i < n; ++i)
            for (int q = 0; q < na; ++q) {
                double t = 0.0;
                for (int c = 0; c < m; ++c) t += x[(size_t)i * m + c] * xtr[c * na + q];
                ra[(size_t)i * na + q] -= t;
            }
        free(xtr);
        if (cholqr(n, ra, na, NULL) != 0) { free(ra); free(ara); break; }
        op_apply(op, ra, na, ara);
        double *pa = NULL, *apa = NULL;
        int use_p = 0;
        if (has_p) {
            pa = (double*)malloc(sizeof(double) * (size_t)n * na);
            apa = (double*)malloc(sizeof(double) * (size_t)n * na);
            for (int64_t i = 0; i < n; ++i) {
                int q = 0;
                for (int c = 0; c < m; ++c)
                    if (act[c]) {
                        pa[(size_t)i * na + q] = p[(size_t)i * m + c];
                        apa[(size_t)i * na + q] = ap[(size_t)i * m + c];
                        ++q;
                    }
            }
            use_p = cholqr(n, pa, na, apa) == 0;
        }
        int sdim = 0, st = -1;
        for (int attempt = 0; attempt < 2 && st != 0; ++attempt) {
            if (attempt == 1) { if (!use_p) break; use_p = 0; }
            sdim = m + na + (use_p ? na : 0);
            for (int64_t i = 0; i < n; ++i) {
                double* si = s + (size_t)i * sdim;
                double* asi = as + (size_t)i * sdim;
                memcpy(si, x + (size_t)i * m, sizeof(double) * m);
                memcpy(asi, ax + (size_t)i * m, sizeof(double) * m);
                memcpy(si + m, ra + (size_t)i * na, sizeof(double) * na);
                memcpy(asi + m, ara + (size_t)i * na, sizeof(double) * na);
                if (use_p) {
                    memcpy(si + m + na, pa + (size_t)i * na, sizeof(double) * na);
                    memcpy(asi + m + na, apa + (size_t)i * na, sizeof(double) * na);
                }
            }
            gram(n, s, sdim, as, sdim, ga);
            gram(n, s, sdim, s, sdim, gb);
            for (int i = 0; i < sdim; ++i)
                for (int j = 0; j < i; ++j) {
                    double a1 = 0.5 * (ga[i * sdim + j] + ga[j * sdim + i]);
                    ga[i * sdim + j] = ga[j * sdim + i] = a1;
                    double b1 = 0.5 * (gb[i * sdim + j] + gb[j * sdim + i]);
                    gb[i * sdim + j] = gb[j * sdim + i] = b1;
                }
            st = or_gen_sym_eig(sdim, ga, gb, th, cc);
        }
        if (st != 0) { free(ra); free(ara); free(pa); free(apa); break; }
        for (int i = 0; i < sdim; ++i)
            for (int c = 0; c < m; ++c) csub[i * m + c] = cc[i * sdim + c];
        /* X = S C, AX = AS C, P = S[:, m:] C[m:], AP likewise */
        for (int64_t i = 0; i < n; ++i) {
            const double* si = s + (size_t)i * sdim;
            const double* asi = as + (size_t)i * sdim;
            for (int c = 0; c < m; ++c) {
                double px = 0.0, pax = 0.0, pp = 0.0, pap = 0.0;
                for (int q = 0; q < m; ++q) { px += si[q] * csub[q * m + c]; pax += asi[q] * csub[q * m + c]; }
                for (int q = m; q < sdim; ++q) { pp += si[q] * csub[q * m + c]; pap += asi[q] * csub[q * m + c]; }
                x[(size_t)i * m + c] = px + pp;
                ax[(size_t)i * m + c] = pax + pap;
                p[(size_t)i * m + c] = pp;
                ap[(size_t)i * m + c] = pap;
            }
        }
        for (int c = 0; c < m; ++c) lam[c] = th[c];
        has_p = 1;
        ++it;
        free(ra); free(ara); free(pa); free(apa);
    }
done:
    if (iters) *iters += it;
    free(ax); free(r); free(p); free(ap); free(s); free(as); free(ga); free(gb);
    free(th); free(cc); free(csub); free(xr); free(act);
    return converged;
}

int or_lobpcg_core(int64_t n, const int64_t* indptr, const int32_t* indices,
                   const double* data, double shift, double* x, int m, double tol,
                   int maxiter, double* evals, double* rnorm, int64_t* iters,
                   int64_t* spmm_calls) {
    op_t op = {n, indptr, indices, data, shift, 0};
    int64_t it = 0;
    int c = lobpcg_core_op(&op, x, m, tol, maxiter, evals, rnorm, &it);
    if (iters) *iters = it;
    if (spmm_calls) *spmm_calls = op.calls;
    return c;
}

/* lobpcg_solve (src/solvers/lobpcg.rs:227-465). */
int or_lobpcg_solve(int64_t n, const int64_t* indptr, const int32_t* indices,
                    const double* data, int n_components, uint64_t seed, int regularize,
                    const double* sqrt_deg, double tol, const double* x0_or_null,
                    double* evals, double* evecs, or_lobpcg_stats* st) {
    const int k = n_components + 1;
    or_lobpcg_stats local = {0, 0, 0, 0, 0};
    if (!st) st = &local;
    memset(st, 0, sizeof(*st));
    if (n_components == 0 || k >= n) return 0; /* lobpcg.rs:239-241 */
    size_t nk = (size_t)n * k;
    op_t op = {n, indptr, indices, data, regularize ? 1e-5 : 0.0, 0};
    double* x = (double*)malloc(sizeof(double) * nk);
    if (x0_or_null) {
        memcpy(x, x0_or_null, sizeof(double) * nk);
    } else {
        /* N(0,1) block (lobpcg.rs:244-246) + trivial hint in column 0 (:249-252) */
        or_fill_normal(seed, STREAM_LOBPCG_INIT, nk, x);
        double sn = 0.0;
        for (int64_t i = 0; i < n; ++i) sn += sqrt_deg[i] * sqrt_deg[i];
        sn = sqrt(sn);
        if (sn > 0.0)
            for (int64_t i = 0; i < n; ++i) x[(size_t)i * k] = sqrt_deg[i] / sn;
        if (n >= 1000) { /* ChFSI (lobpcg.rs:270-291) */
            double* t = (double*)malloc(sizeof(double) * nk);
            or_chebyshev_filter(n, indptr, indices, data, x, k, 0.1, 2.0, 8, t);
            op.calls += 8;
            for (size_t i = 0; i < nk; ++i) t[i] = (x[i] + t[i]) / 2.0;
            for (int c = 0; c < k; ++c) {
                double q = 0.0;
                for (int64_t i = 0; i < n; ++i) q += t[(size_t)i * k + c] * t[(size_t)i * k + c];
                q = sqrt(q);
                for (int64_t i = 0; i < n; ++i)
                    x[(size_t)i * k + c] = q > 1e-300 ? t[(size_t)i * k + c] / q : t[(size_t)i * k + c];
            }
            free(t);
        }
    }
    double* lam = (double*)malloc(sizeof(double) * k);
    double* rn = (double*)malloc(sizeof(double) * k);
    double* res = (double*)malloc(sizeof(double) * k);
    double* best = NULL;
    double* best_l = (double*)malloc(sizeof(double) * k);
    int have_last = 0, produced = 0;
    int restart_count = 0;
    for (int restart = 0; restart <= 3; ++restart) {
        int maxiter = (int)((5 * n) < (restart == 0 ? 300 : 1000) ? 5 * n : (restart == 0 ? 300 : 1000));
        int conv = lobpcg_core_op(&op, x, k, tol, maxiter, lam, rn, &st->iterations);
        st->passes++;
        int accept = conv;
        if (!conv) {
            accept = 1;
            for (int c = 0; c < k; ++c) if (!(rn[c] < tol)) accept = 0;
        }
        if (!accept) {
            if (restart < 3) continue; /* x already holds the partial eigvecs */
            break;
        }
        /* Rayleigh-Ritz refinement against the true Laplacian */
        double* xr = (double*)malloc(sizeof(double) * nk);
        memcpy(xr, x, sizeof(double) * nk);
        if (rr_refine(&op, xr, k, lam) != 0) {
            free(xr);
            if (restart < 3) continue;
            break;
        }
        or_per_vector_residuals(n, indptr, indices, data, lam, xr, k, res);
        op.calls++;
        int all_ok = 1;
        for (int c = 0; c < k; ++c) if (!(res[c] < tol)) all_ok = 0;
        if (all_ok) {
            memcpy(evecs, xr, sizeof(double) * nk);
            memcpy(evals, lam, sizeof(double) * k);
            free(xr);
            produced = 1;
            st->converged = 1;
            st->restarts = restart_count;
            goto out;
        }
        restart_count++;
        if (!best) best = (double*)malloc(sizeof(double) * nk);
        memcpy(best, xr, sizeof(double) * nk);
        memcpy(best_l, lam, sizeof(double) * k);
        have_last = 1;
        memcpy(x, xr, sizeof(double) * nk); /* warm restart from refined vectors */
        free(xr);
    }
    if (have_last) {
        memcpy(evecs, best, sizeof(double) * nk);
        memcpy(evals, best_l, sizeof(double) * k);
        produced = 1;
        st->restarts = restart_count;
    }
out:
    st->spmm_calls = op.calls;
    free(x); free(lam); free(rn); free(res); free(best); free(best_l);
    return produced;
}

/* ------------------------------------------------------------------------ */
/* Thin Q of a QR factorisation by twice-repeated modified Gram-Schmidt
 * (same Q as Householder QR up to column signs). */
static void qr_thin_q(int64_t n, double* y, int k) {
    double* col = (double*)malloc(sizeof(double) * (size_t)n);
    for (int j = 0; j < k; ++j) {
        for (int64_t i = 0; i < n; ++i) col[i] = y[(size_t)i * k + j];
        for (int pass = 0; pass < 2; ++pass)
            for (int q = 0; q < j; ++q) {
                double d = 0.0;
                for (int64_t i = 0; i < n; ++i) d += y[(size_t)i * k + q] * col[i];
                for (int64_t i = 0; i < n; ++i) col[i] -= d * y[(size_t)i * k + q];
            }
        double nr = 0.0;
        for (int64_t i = 0; i < n; ++i) nr += col[i] * col[i];
        nr = sqrt(nr);
        for (int64_t i = 0; i < n; ++i) y[(size_t)i * k + j] = nr > 0 ? col[i] / nr : 0.0;
    }
    free(col);
}

/* rsvd_solve (src/solvers/rsvd.rs:135-213). */
void or_rsvd_solve(int64_t n, const int64_t* indptr, const int32_t* indices,
                   const double* data, int n_components, uint64_t seed, int oversampling,
                   int nbiter, double* evals, double* evecs) {
    int rank = n_components + 1;
    int64_t ksub;
    if (oversampling < 0) { /* rsvd_k_sub (rsvd.rs:84-87) */
        int64_t os = n / 10;
        int64_t mr = rank > 5 ? rank : 5;
        if (os < mr) os = mr;
        int64_t cap = n > rank ? n - rank : 0;
        if (os > cap) os = cap;
        ksub = rank + os;
    } else {
        ksub = rank + oversampling;
    }
    if (ksub > n) ksub = n;
    int k = (int)ksub;
    /* M = 2I - L (rsvd.rs:16-28) */
    int64_t nnz = indptr[n];
    double* md = (double*)malloc(sizeof(double) * (size_t)nnz);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t p = indptr[i]; p < indptr[i + 1]; ++p)
            md[p] = indices[p] == i ? 2.0 - data[p] : -data[p];
    size_t nk = (size_t)n * k;
    double* om = (double*)malloc(sizeof(double) * nk);
    double* y = (double*)malloc(sizeof(double) * nk);
    or_fill_normal(seed, STREAM_RSVD_OMEGA, nk, om);
    for (int it = 0; it < nbiter; ++it) {
        or_spmm(n, indptr, indices, md, om, k, y);
        qr_thin_q(n, y, k);
        or_spmm(n, indptr, indices, md, y, k, om);
        qr_thin_q(n, om, k);
    }
    or_spmm(n, indptr, indices, md, om, k, y);
    qr_thin_q(n, y, k); /* Q */
    or_spmm(n, indptr, indices, md, y, k, om); /* MQ */
    double* b = (double*)malloc(sizeof(double) * (size_t)k * k);
    gram(n, y, k, om, k, b);
    double* w = (double*)malloc(sizeof(double) * (size_t)k);
    double* u = (double*)malloc(sizeof(double) * (size_t)k * k);
    or_sym_eig(k, b, w, u);
    block_mul(n, y, k, u, k, om); /* V = Q U */
    int idx_trivial = k - 1;
    evals[0] = 2.0 - w[idx_trivial];
    for (int64_t i = 0; i < n; ++i) evecs[(size_t)i * rank] = om[(size_t)i * k + idx_trivial];
    for (int c = 0; c < n_components; ++c) {
        int idx = k - 2 - c;
        if (idx < 0) idx = 0;
        evals[c + 1] = 2.0 - w[idx];
        for (int64_t i = 0; i < n; ++i) evecs[(size_t)i * rank + c + 1] = om[(size_t)i * k + idx];
    }
    free(md); free(om); free(y); free(b); free(w); free(u);
}

/* solve_eigenproblem (src/solvers/mod.rs:70-224); thresholds metrics.rs:17-46. */
int or_solve_eigenproblem(int64_t n, const int64_t* indptr, const int32_t* indices,
                          const double* data, int n_components, uint64_t seed,
                          const double* sqrt_deg, int64_t dense_threshold, double lobpcg_tol,
                          double* evals, double* evecs, or_lobpcg_stats* st) {
    int k = n_components + 1;
    if (dense_threshold < 0) dense_threshold = 2000;
    if (lobpcg_tol <= 0) lobpcg_tol = 1e-5;
    /* LOBPCG_QUALITY_THRESHOLD equals LOBPCG_ACCEPT_TOL in the reference (1e-5);
     * a tighter solver tolerance tightens the gate with it. */
    double gate_lobpcg = lobpcg_tol;
    if (n < dense_threshold) {
        if (or_dense_evd(n, indptr, indices, data, k, evals, evecs) == 0) {
            double q = or_max_eigenpair_residual(n, indptr, indices, data, evecs, evals, k);
            if (q < 1e-6) return 0;
        }
    }
    if (or_lobpcg_solve(n, indptr, indices, data, n_components, seed, 0, sqrt_deg, lobpcg_tol,
                        NULL, evals, evecs, st)) {
        double q = or_max_eigenpair_residual(n, indptr, indices, data, evecs, evals, k);
        if (q < gate_lobpcg) return 1;
    }
    /* Level 2 (shift-invert via sparse Cholesky) is not restated here. */
    if (or_lobpcg_solve(n, indptr, indices, data, n_components, seed, 1, sqrt_deg, lobpcg_tol,
                        NULL, evals, evecs, st)) {
        double q = or_max_eigenpair_residual(n, indptr, indices, data, evecs, evals, k);
        if (q < gate_lobpcg) return 3;
    }
    or_rsvd_solve(n, indptr, indices, data, n_components, seed, -1, 2, evals, evecs);
    {
        double q = or_max_eigenpair_residual(n, indptr, indices, data, evecs, evals, k);
        if (q < 1e-2) return 4;
    }
    or_dense_evd(n, indptr, indices, data, k, evals, evecs);
    return 5;
}
