/*
 * oracle_graph.c -- CPU restatement of the fuzzy k-NN graph that feeds the
 * eigensolver path (the input W of spectral_init).  TEST INFRASTRUCTURE ONLY:
 * the checker for the device graph builder and the CPU baseline beside it.
 *
 * The reference builds its fixture graphs in tests/generate_fixtures.py with
 * umap-learn 0.5.11 (meta.json of every fixture; umap is not installed here):
 *   step 1  exact k-NN, self excluded            generate_fixtures.py:70-93
 *   step 2  umap smooth_knn_dist (sigma, rho)    generate_fixtures.py:96-111
 *   step 3  umap compute_membership_strengths    generate_fixtures.py:114-141
 *   step 4  fuzzy union A + A^T - A o A^T        generate_fixtures.py:144-161
 *   step 5a prune below max / n_epochs           generate_fixtures.py:164-182
 * Steps 2 and 3 restate umap-learn 0.5.11's published algorithms
 * (umap/umap_.py: smooth_knn_dist with n_iter 64, local_connectivity 1,
 * bandwidth 1, SMOOTH_K_TOLERANCE 1e-5, MIN_K_DIST_SCALE 1e-3, and psum / lo /
 * mid / hi held in float32; compute_membership_strengths in float32).
 *
 * Pinned by the reference fixtures step1_knn_exact .. step5a_pruned_exact of
 * the nine datasets (each step fed the fixture's own previous step).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* Exact k-NN by brute force, self excluded (generate_fixtures.py:78-86:
 * pairwise_distances, diagonal set to inf, argsort per row).  Distances are
 * FP64 sums of squared differences in dimension order, each term fused
 * (fma), sqrt'ed and rounded to float32; ties go to the smaller index. */
void or_knn_exact(int64_t n, int d, const double* x, int k, int32_t* idx, float* dist) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        double* bd = (double*)malloc(sizeof(double) * (size_t)k);
        int32_t* bi = (int32_t*)malloc(sizeof(int32_t) * (size_t)k);
        int cnt = 0;
        const double* xi = x + i * d;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            const double* xj = x + j * d;
            double s = 0.0;
            for (int t = 0; t < d; ++t) {
                const double df = xi[t] - xj[t];
                s = fma(df, df, s);
            }
            if (cnt == k && !(s < bd[k - 1])) continue;
            int p = cnt < k ? cnt++ : k - 1;
            while (p > 0 && s < bd[p - 1]) {
                bd[p] = bd[p - 1];
                bi[p] = bi[p - 1];
                --p;
            }
            bd[p] = s;
            bi[p] = (int32_t)j;
        }
        for (int t = 0; t < k; ++t) {
            idx[i * k + t] = t < cnt ? bi[t] : -1;
            dist[i * k + t] = t < cnt ? (float)sqrt(bd[t]) : INFINITY;
        }
        free(bd);
        free(bi);
    }
}

/* umap smooth_knn_dist (umap 0.5.11), called as smooth_knn_dist(knn_dists, k)
 * by generate_fixtures.py:103.  Per row: rho = first non-zero distance; sigma
 * by bisection so that sum_{j>=1} exp(-max(d_j - rho, 0) / sigma) = log2(k)
 * (a term with d_j - rho <= 0 counts 1), at most 64 steps, stopping within
 * 1e-5; then sigma >= 1e-3 * mean(row distances) when rho > 0, else
 * >= 1e-3 * mean(all distances).  Arithmetic in float32 as umap declares it.
 *
 * umap compiles this function with fastmath (no-infs), which folds its
 * "hi == inf" test to false: where the upper bracket is still open, the step
 * that should double mid computes (lo + inf) / 2 = inf instead, and the row
 * ends with sigma = inf (every membership of that row is then 1).  The
 * fixtures hold exactly that (all 1952 such rows of the nine datasets), so it
 * is restated here, together with fastmath's reciprocal form d * (1 / mid). */
void or_smooth_knn_dist(int64_t n, int k, const float* dist, float* sigmas, float* rhos) {
    const float target = (float)(log2((double)k));
    /* mean of all distances, used only by rows without a non-zero distance;
     * accumulated in double (no fixture row exercises it) */
    double sum_all = 0.0;
    for (int64_t t = 0; t < n * (int64_t)k; ++t) sum_all += (double)dist[t];
    const float mean_all = (float)(sum_all / (double)(n * (int64_t)k));
    for (int64_t i = 0; i < n; ++i) {
        const float* di = dist + i * k;
        float rho = 0.0f;
        for (int j = 0; j < k; ++j)
            if (di[j] > 0.0f) { rho = di[j]; break; }
        float lo = 0.0f, hi = INFINITY, mid = 1.0f;
        for (int it = 0; it < 64; ++it) {
            float psum = 0.0f;
            for (int j = 1; j < k; ++j) {
                const float dd = di[j] - rho;
                psum += dd > 0.0f ? expf(-dd * (1.0f / mid)) : 1.0f;
            }
            if (fabsf(psum - target) < 1e-5f) break;
            if (psum > target) hi = mid;
            else lo = mid;
            mid = (lo + hi) / 2.0f;  /* inf while hi is open (see above) */
        }
        float sigma = mid;
        if (rho > 0.0f) {
            float m = 0.0f;
            for (int j = 0; j < k; ++j) m += di[j];
            m /= (float)k;
            if ((double)sigma < 1e-3 * (double)m) sigma = (float)(1e-3 * (double)m);
        } else if ((double)sigma < 1e-3 * (double)mean_all) {
            sigma = (float)(1e-3 * (double)mean_all);
        }
        sigmas[i] = sigma;
        rhos[i] = rho;
    }
}

/* umap compute_membership_strengths (umap 0.5.11), generate_fixtures.py:127-132:
 * val = 0 for a self entry, 1 when d - rho <= 0 or sigma == 0, otherwise
 * exp(-(d - rho) / sigma), all in float32; index -1 (missing) gives 0. */
void or_membership(int64_t n, int k, const int32_t* idx, const float* dist, const float* sigmas,
                   const float* rhos, float* vals) {
    for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < k; ++j) {
            const int64_t p = i * k + j;
            float v;
            if (idx[p] < 0 || idx[p] == i) v = 0.0f;
            else if (dist[p] - rhos[i] <= 0.0f || sigmas[i] == 0.0f) v = 1.0f;
            else v = expf(-((dist[p] - rhos[i]) / sigmas[i]));
            vals[p] = v;
        }
}

typedef struct {
    int64_t key;  /* row * n + col */
    float a;      /* a_rc (entry of A) */
    float at;     /* a_cr (entry of A^T) */
} UEnt;

static int cmp_uent(const void* pa, const void* pb) {
    const int64_t a = ((const UEnt*)pa)->key, b = ((const UEnt*)pb)->key;
    return a < b ? -1 : (a > b ? 1 : 0);
}

/* Fuzzy union of the directed membership graph A (row i holds the k entries
 * (i, idx[i,j], vals[i,j]); zeros dropped first, generate_fixtures.py:138):
 * W = (A + A^T) - A o A^T in float32, zeros eliminated, rows sorted by column
 * (generate_fixtures.py:155-158).  Then, when n_epochs > 0, entries below
 * max(W) / n_epochs (a float32 quotient) are dropped (generate_fixtures.py:
 * 176-179).  Returns nnz; with indices == NULL only counts. */
int64_t or_fuzzy_union(int64_t n, int k, const int32_t* idx, const float* vals, int n_epochs,
                       int64_t* indptr, int32_t* indices, float* data) {
    const int64_t m = n * (int64_t)k;
    UEnt* e = (UEnt*)malloc(sizeof(UEnt) * (size_t)(2 * m + 1));
    int64_t c = 0;
    for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < k; ++j) {
            const int64_t p = i * k + j;
            if (idx[p] < 0 || vals[p] == 0.0f) continue;
            const int64_t col = idx[p];
            e[c].key = i * n + col; e[c].a = vals[p]; e[c].at = 0.0f; ++c;
            e[c].key = col * n + i; e[c].a = 0.0f; e[c].at = vals[p]; ++c;
        }
    qsort(e, (size_t)c, sizeof(UEnt), cmp_uent);
    /* merge equal keys: at most one A entry and one A^T entry per key */
    int64_t u = 0;
    for (int64_t t = 0; t < c; ++t) {
        if (u > 0 && e[u - 1].key == e[t].key) {
            e[u - 1].a += e[t].a;
            e[u - 1].at += e[t].at;
        } else {
            e[u++] = e[t];
        }
    }
    float wmax = 0.0f;
    for (int64_t t = 0; t < u; ++t) {
        const float s = e[t].a + e[t].at;
        const float pr = e[t].a * e[t].at;
        e[t].a = s - pr;
        if (e[t].a > wmax) wmax = e[t].a;
    }
    const float thr = n_epochs > 0 ? wmax / (float)n_epochs : 0.0f;
    int64_t nnz = 0;
    if (indptr) for (int64_t i = 0; i <= n; ++i) indptr[i] = 0;
    for (int64_t t = 0; t < u; ++t) {
        const float w = e[t].a;
        if (w == 0.0f || (n_epochs > 0 && w < thr)) continue;
        if (indices) {
            indices[nnz] = (int32_t)(e[t].key % n);
            data[nnz] = w;
        }
        if (indptr) indptr[e[t].key / n + 1] += 1;
        ++nnz;
    }
    if (indptr)
        for (int64_t i = 0; i < n; ++i) indptr[i + 1] += indptr[i];
    free(e);
    return nnz;
}
