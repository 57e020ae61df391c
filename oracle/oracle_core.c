/*
 * oracle_core.c -- counter RNG, degrees, Laplacian, CSR SpMM, residuals and
 * the small/dense FP64 eigen kernels of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Reference files
 * are cited per function; the dense kernels restate textbook methods
 * (Cholesky, cyclic Jacobi, Householder tridiagonalisation + implicit QL) in
 * place of faer's SelfAdjointEigen, which is not vendored in the reference.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;

void or_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int or_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11) and a Box-Muller transform.  The device path evaluates the same
 * function, so oracle and GPU solves can start from identical blocks.        */
static void philox4x32_10(uint32_t ctr[4], uint32_t key0, uint32_t key1) {
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * ctr[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * ctr[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ ctr[1] ^ key0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ ctr[3] ^ key1;
        uint32_t n3 = (uint32_t)p0;
        ctr[0] = n0; ctr[1] = n1; ctr[2] = n2; ctr[3] = n3;
        key0 += 0x9E3779B9u;
        key1 += 0xBB67AE85u;
    }
}

/* The raw generator on one counter block (the known-answer vectors of the
 * published Random123 suite check it). */
void or_philox4x32_10(const uint32_t ctr[4], uint32_t key0, uint32_t key1, uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox4x32_10(c, key0, key1);
    for (int i = 0; i < 4; ++i) out[i] = c[i];
}

double or_normal(uint64_t seed, uint32_t stream, uint64_t index) {
    uint32_t c[4] = {(uint32_t)index, (uint32_t)(index >> 32), stream, 0x5eedu};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double inv53 = 1.0 / 9007199254740992.0; /* 2^-53 */
    uint64_t a = ((uint64_t)c[0] << 21) ^ (uint64_t)(c[1] >> 11);
    uint64_t b = ((uint64_t)c[2] << 21) ^ (uint64_t)(c[3] >> 11);
    a &= (1ull << 53) - 1;
    b &= (1ull << 53) - 1;
    double u1 = ((double)a + 1.0) * inv53; /* (0, 1] */
    double u2 = (double)b * inv53;         /* [0, 1) */
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
}

void or_fill_normal(uint64_t seed, uint32_t stream, size_t count, double* out) {
    for (size_t i = 0; i < count; ++i) out[i] = or_normal(seed, stream, (uint64_t)i);
}

/* The FP32 start-block stream of the device's FP32-bulk solvers: pair p of
 * the counter space gives the cosine and sine outputs of one Box-Muller
 * transform in float, from 24-bit uniforms of the first two Philox words. */
void or_normal_pairs_f32(uint64_t seed, uint32_t stream, size_t npairs, float* out) {
    const float inv24 = 1.0f / 16777216.0f; /* 2^-24 */
    for (size_t p = 0; p < npairs; ++p) {
        uint32_t c[4] = {(uint32_t)p, (uint32_t)((uint64_t)p >> 32), stream, 0x9a1fu};
        philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        const float u1 = ((float)(c[0] >> 8) + 1.0f) * inv24; /* (0, 1] */
        const float u2 = (float)(c[1] >> 8) * inv24;          /* [0, 1) */
        const float r = sqrtf(-2.0f * logf(u1));
        const float t = 6.2831853071795865f * u2;
        out[2 * p] = r * cosf(t);
        out[2 * p + 1] = r * sinf(t);
    }
}

/* ------------------------------------------------------------------------ */
/* compute_degrees (src/laplacian.rs:29-49).                                  */
void or_compute_degrees(int64_t n, const int64_t* indptr, const int32_t* indices,
                        const float* data, int mode, double* deg, double* sqrt_deg) {
    if (mode == 0) {
        /* column_sum_f32 (laplacian.rs:8-16): deg_f32[col] += val in CSR order. */
        float* acc = (float*)calloc((size_t)n, sizeof(float));
        for (int64_t i = 0; i < n; ++i)
            for (int64_t p = indptr[i]; p < indptr[i + 1]; ++p) acc[indices[p]] += data[p];
        for (int64_t i = 0; i < n; ++i) deg[i] = (double)acc[i];
        free(acc);
    } else {
        /* RustNative: f64 row sums (laplacian.rs:41-46). */
        for (int64_t i = 0; i < n; ++i) {
            double s = 0.0;
            for (int64_t p = indptr[i]; p < indptr[i + 1]; ++p) s += (double)data[p];
            deg[i] = s;
        }
    }
    for (int64_t i = 0; i < n; ++i) sqrt_deg[i] = sqrt(deg[i]);
}

int64_t or_laplacian_nnz(int64_t n, const int64_t* indptr, const int32_t* indices) {
    int64_t nnz = n;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t p = indptr[i]; p < indptr[i + 1]; ++p)
            if (indices[p] != i) ++nnz;
    return nnz; /* upper bound; duplicate (i,j) pairs are merged below */
}

/* build_normalized_laplacian (src/laplacian.rs:60-88): off-diagonal
 * -isd[i]*w*isd[j] evaluated left to right, unit diagonal for every row,
 * TriMat::to_csr ordering (columns ascending, duplicates summed).             */
void or_build_laplacian(int64_t n, const int64_t* indptr, const int32_t* indices,
                        const float* data, const double* inv_sqrt_deg,
                        int64_t* l_indptr, int32_t* l_indices, double* l_data) {
    int64_t out = 0;
    l_indptr[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        int64_t row_start = out;
        for (int64_t p = indptr[i]; p < indptr[i + 1]; ++p) {
            int32_t j = indices[p];
            if (j == i) continue;
            l_indices[out] = j;
            l_data[out] = -inv_sqrt_deg[i] * (double)data[p] * inv_sqrt_deg[j];
            ++out;
        }
        l_indices[out] = (int32_t)i;
        l_data[out] = 1.0;
        ++out;
        /* stable insertion sort by column (rows are short and nearly sorted) */
        for (int64_t a = row_start + 1; a < out; ++a) {
            int32_t cj = l_indices[a];
            double cv = l_data[a];
            int64_t b = a - 1;
            while (b >= row_start && l_indices[b] > cj) {
                l_indices[b + 1] = l_indices[b];
                l_data[b + 1] = l_data[b];
                --b;
            }
            l_indices[b + 1] = cj;
            l_data[b + 1] = cv;
        }
        /* merge duplicates by summation */
        int64_t w = row_start;
        for (int64_t a = row_start; a < out; ++a) {
            if (a > row_start && l_indices[a] == l_indices[w - 1]) {
                l_data[w - 1] += l_data[a];
            } else {
                l_indices[w] = l_indices[a];
                l_data[w] = l_data[a];
                ++w;
            }
        }
        out = w;
        l_indptr[i + 1] = out;
    }
}

/* ------------------------------------------------------------------------ */
/* CsrOperator::apply (src/operator.rs:176-228): zero-filled y, then for each
 * row the stored entries in CSR order accumulate into y[i, :] (the order of
 * sprs::prod::csr_mulacc_dense_rowmaj and of spmv_csr for k = 1).            */
void or_spmm(int64_t n, const int64_t* indptr, const int32_t* indices, const double* data,
             const double* x, int k, double* y) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
#endif
    for (int64_t i = 0; i < n; ++i) {
        double* yi = y + (size_t)i * k;
        for (int c = 0; c < k; ++c) yi[c] = 0.0;
        for (int64_t p = indptr[i]; p < indptr[i + 1]; ++p) {
            const double v = data[p];
            const double* xj = x + (size_t)indices[p] * k;
            for (int c = 0; c < k; ++c) yi[c] += v * xj[c];
        }
    }
}

/* eigenpair_residual (src/metrics.rs:57-68). */
double or_eigenpair_residual(int64_t n, const int64_t* indptr, const int32_t* indices,
                             const double* data, const double* v, double lambda) {
    double num = 0.0, den = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int64_t p = indptr[i]; p < indptr[i + 1]; ++p) acc += data[p] * v[indices[p]];
        double d = acc - lambda * v[i];
        num += d * d;
        den += v[i] * v[i];
    }
    den = sqrt(den);
    if (den < 1e-300) den = 1e-300;
    return sqrt(num) / den;
}

/* max_eigenpair_residual (src/metrics.rs:73-92): NaN-propagating max. */
double or_max_eigenpair_residual(int64_t n, const int64_t* indptr, const int32_t* indices,
                                 const double* data, const double* vecs, const double* lambdas,
                                 int k) {
    double* col = (double*)malloc(sizeof(double) * (size_t)n);
    double worst = 0.0;
    for (int c = 0; c < k; ++c) {
        for (int64_t i = 0; i < n; ++i) col[i] = vecs[(size_t)i * k + c];
        double r = or_eigenpair_residual(n, indptr, indices, data, col, lambdas[c]);
        if (isnan(worst) || isnan(r)) worst = NAN;
        else if (r > worst) worst = r;
    }
    free(col);
    return worst;
}

/* ------------------------------------------------------------------------ */
/* Small dense FP64 kernels.                                                   */

/* In-place lower Cholesky A = L L^T (row-major).  Returns 0 on success, -1 if
 * a pivot is not positive. The strict upper triangle is zeroed. */
int or_cholesky(int n, double* a) {
    for (int j = 0; j < n; ++j) {
        double s = a[j * n + j];
        for (int p = 0; p < j; ++p) s -= a[j * n + p] * a[j * n + p];
        if (!(s > 0.0)) return -1;
        double d = sqrt(s);
        a[j * n + j] = d;
        for (int i = j + 1; i < n; ++i) {
            double t = a[i * n + j];
            for (int p = 0; p < j; ++p) t -= a[i * n + p] * a[j * n + p];
            a[i * n + j] = t / d;
        }
        for (int i = 0; i < j; ++i) a[i * n + j] = 0.0;
    }
    return 0;
}

static void sort_eig_ascending(int n, double* w, double* v /* n x n, column j = vector j */) {
    for (int i = 0; i < n; ++i) {
        int best = i;
        for (int j = i + 1; j < n; ++j)
            if (w[j] < w[best]) best = j;
        if (best != i) {
            double t = w[i]; w[i] = w[best]; w[best] = t;
            for (int r = 0; r < n; ++r) {
                double u = v[r * n + i]; v[r * n + i] = v[r * n + best]; v[r * n + best] = u;
            }
        }
    }
}

/* Cyclic Jacobi eigen-decomposition of a symmetric matrix (the lower triangle
 * of a is read, matching SelfAdjointEigen(Side::Lower)).  w ascending,
 * v column j = eigenvector j. */
void or_sym_eig(int n, const double* a_in, double* w, double* v) {
    double* a = (double*)malloc(sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) a[i * n + j] = a[j * n + i] = a_in[i * n + j];
    for (int i = 0; i < n * n; ++i) v[i] = 0.0;
    for (int i = 0; i < n; ++i) v[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double q = a[i * n + j] * a[i * n + j];
                tot += q;
                if (i != j) off += q;
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = a[p * n + q];
                if (apq == 0.0) continue;
                double app = a[p * n + p], aqq = a[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int r = 0; r < n; ++r) {
                    double arp = a[r * n + p], arq = a[r * n + q];
                    a[r * n + p] = c * arp - s * arq;
                    a[r * n + q] = s * arp + c * arq;
                }
                for (int r = 0; r < n; ++r) {
                    double apr = a[p * n + r], aqr = a[q * n + r];
                    a[p * n + r] = c * apr - s * aqr;
                    a[q * n + r] = s * apr + c * aqr;
                }
                a[p * n + q] = a[q * n + p] = 0.0;
                for (int r = 0; r < n; ++r) {
                    double vrp = v[r * n + p], vrq = v[r * n + q];
                    v[r * n + p] = c * vrp - s * vrq;
                    v[r * n + q] = s * vrp + c * vrq;
                }
            }
    }
    for (int i = 0; i < n; ++i) w[i] = a[i * n + i];
    sort_eig_ascending(n, w, v);
    free(a);
}

/* Generalised symmetric-definite problem A c = theta B c via B = L L^T,
 * M = L^-1 A L^-T, c = L^-T y.  Returns -1 if B is not positive definite. */
int or_gen_sym_eig(int n, const double* a, const double* b, double* w, double* c) {
    size_t nn = (size_t)n * n;
    double* l = (double*)malloc(sizeof(double) * nn);
    double* m = (double*)malloc(sizeof(double) * nn);
    double* y = (double*)malloc(sizeof(double) * nn);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) l[i * n + j] = l[j * n + i] = b[i * n + j];
    if (or_cholesky(n, l) != 0) { free(l); free(m); free(y); return -1; }
    /* m = L^-1 A (forward substitution per column), then m = (L^-1 m^T)^T */
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) m[i * n + j] = m[j * n + i] = a[i * n + j];
    for (int col = 0; col < n; ++col)
        for (int i = 0; i < n; ++i) {
            double s = m[i * n + col];
            for (int p = 0; p < i; ++p) s -= l[i * n + p] * m[p * n + col];
            m[i * n + col] = s / l[i * n + i];
        }
    /* now m = L^-1 A ; compute L^-1 m^T and transpose */
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) y[i * n + j] = m[j * n + i];
    for (int col = 0; col < n; ++col)
        for (int i = 0; i < n; ++i) {
            double s = y[i * n + col];
            for (int p = 0; p < i; ++p) s -= l[i * n + p] * y[p * n + col];
            y[i * n + col] = s / l[i * n + i];
        }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) m[i * n + j] = 0.5 * (y[i * n + j] + y[j * n + i]);
    or_sym_eig(n, m, w, y);
    /* c = L^-T y: back substitution per column */
    for (int col = 0; col < n; ++col)
        for (int i = n - 1; i >= 0; --i) {
            double s = y[i * n + col];
            for (int p = i + 1; p < n; ++p) s -= l[p * n + i] * c[p * n + col];
            c[i * n + col] = s / l[i * n + i];
        }
    free(l); free(m); free(y);
    return 0;
}

/* Householder tridiagonalisation + implicit-shift QL for all eigenpairs of a
 * dense symmetric matrix (full storage, row-major).  zt row i = eigenvector i.
 * Returns 0 on success, -1 if QL does not converge. */
static int dense_sym_eig_all(int n, double* a, double* d, double* zt) {
    double* e = (double*)calloc((size_t)n, sizeof(double));
    double* vs = (double*)calloc((size_t)n * n, sizeof(double)); /* Householder vectors */
    double* taus = (double*)calloc((size_t)n, sizeof(double));
    double* p = (double*)malloc(sizeof(double) * (size_t)n);
    for (int k = 0; k + 2 < n; ++k) {
        int m = n - k - 1;
        double* v = vs + (size_t)k * n; /* v[0..m) for rows k+1..n */
        double xn = 0.0;
        for (int r = 0; r < m; ++r) { v[r] = a[(size_t)(k + 1 + r) * n + k]; xn += v[r] * v[r]; }
        xn = sqrt(xn);
        if (xn == 0.0) { taus[k] = 0.0; continue; }
        double alpha = v[0] > 0 ? -xn : xn;
        v[0] -= alpha;
        double vn2 = 0.0;
        for (int r = 0; r < m; ++r) vn2 += v[r] * v[r];
        if (vn2 == 0.0) { taus[k] = 0.0; continue; }
        double tau = 2.0 / vn2;
        taus[k] = tau;
        /* p = tau * A22 v */
        for (int r = 0; r < m; ++r) {
            const double* row = a + (size_t)(k + 1 + r) * n + (k + 1);
            double s = 0.0;
            for (int c = 0; c < m; ++c) s += row[c] * v[c];
            p[r] = tau * s;
        }
        double kk = 0.0;
        for (int r = 0; r < m; ++r) kk += v[r] * p[r];
        kk *= 0.5 * tau;
        for (int r = 0; r < m; ++r) p[r] -= kk * v[r];
        for (int r = 0; r < m; ++r) {
            double* row = a + (size_t)(k + 1 + r) * n + (k + 1);
            for (int c = 0; c < m; ++c) row[c] -= v[r] * p[c] + p[r] * v[c];
        }
        a[(size_t)(k + 1) * n + k] = a[(size_t)k * n + k + 1] = alpha;
        for (int r = 1; r < m; ++r) a[(size_t)(k + 1 + r) * n + k] = a[(size_t)k * n + k + 1 + r] = 0.0;
    }
    for (int i = 0; i < n; ++i) d[i] = a[(size_t)i * n + i];
    for (int i = 0; i + 1 < n; ++i) e[i] = a[(size_t)(i + 1) * n + i];
    e[n - 1] = 0.0;
    /* zt = Q^T = H_{n-3} ... H_0 ; build by left-applying H_0 first. */
    memset(zt, 0, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; ++i) zt[(size_t)i * n + i] = 1.0;
    for (int k = 0; k + 2 < n; ++k) {
        double tau = taus[k];
        if (tau == 0.0) continue;
        int m = n - k - 1;
        const double* v = vs + (size_t)k * n;
        for (int c = 0; c < n; ++c) p[c] = 0.0;
        for (int r = 0; r < m; ++r) {
            const double* row = zt + (size_t)(k + 1 + r) * n;
            for (int c = 0; c < n; ++c) p[c] += v[r] * row[c];
        }
        for (int r = 0; r < m; ++r) {
            double* row = zt + (size_t)(k + 1 + r) * n;
            double f = tau * v[r];
            for (int c = 0; c < n; ++c) row[c] -= f * p[c];
        }
    }
    /* implicit QL with Wilkinson-style shifts on (d, e); rotations act on
     * rows i, i+1 of zt (= columns of Q). */
    int status = 0;
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m < n - 1; ++m) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
            }
            if (m == l) break;
            if (++iter > 60) { status = -1; break; }
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
            double s = 1.0, c = 1.0, pp = 0.0;
            int i;
            int underflow = 0;
            for (i = m - 1; i >= l; --i) {
                double f = s * e[i], b = c * e[i];
                r = hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) { d[i + 1] -= pp; e[m] = 0.0; underflow = 1; break; }
                s = f / r;
                c = g / r;
                g = d[i + 1] - pp;
                r = (d[i] - g) * s + 2.0 * c * b;
                pp = s * r;
                d[i + 1] = g + pp;
                g = c * r - b;
                double* zi = zt + (size_t)i * n;
                double* zi1 = zt + (size_t)(i + 1) * n;
                for (int q = 0; q < n; ++q) {
                    double t = zi1[q];
                    zi1[q] = s * zi[q] + c * t;
                    zi[q] = c * zi[q] - s * t;
                }
            }
            if (underflow) continue;
            d[l] -= pp;
            e[l] = g;
            e[m] = 0.0;
        }
        if (status) break;
    }
    free(e); free(vs); free(taus); free(p);
    return status;
}

/* dense_evd (src/solvers/dense.rs:9-46): densify, full symmetric EVD, keep the
 * k smallest eigenpairs in ascending order. */
int or_dense_evd(int64_t n, const int64_t* indptr, const int32_t* indices, const double* data,
                 int k, double* evals, double* evecs) {
    if (k > n) return -1;
    size_t nn = (size_t)n * n;
    double* a = (double*)calloc(nn, sizeof(double));
    double* zt = (double*)malloc(sizeof(double) * nn);
    double* d = (double*)malloc(sizeof(double) * (size_t)n);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t p = indptr[i]; p < indptr[i + 1]; ++p) a[(size_t)i * n + indices[p]] = data[p];
    /* SelfAdjointEigen(Side::Lower): the lower triangle defines the matrix */
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = i + 1; j < n; ++j) a[(size_t)i * n + j] = a[(size_t)j * n + i];
    int st = dense_sym_eig_all((int)n, a, d, zt);
    if (st == 0) {
        int* order = (int*)malloc(sizeof(int) * (size_t)n);
        for (int i = 0; i < n; ++i) order[i] = i;
        for (int i = 1; i < n; ++i) { /* insertion sort of indices by eigenvalue */
            int oi = order[i];
            int j = i - 1;
            while (j >= 0 && d[order[j]] > d[oi]) { order[j + 1] = order[j]; --j; }
            order[j + 1] = oi;
        }
        for (int c = 0; c < k; ++c) {
            evals[c] = d[order[c]];
            const double* z = zt + (size_t)order[c] * n;
            for (int64_t i = 0; i < n; ++i) evecs[(size_t)i * k + c] = z[i];
        }
        free(order);
    }
    free(a); free(zt); free(d);
    return st;
}
