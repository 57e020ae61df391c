/*
 * oracle_pipeline.c -- the spectral_init pipeline around the eigensolver:
 * connected components, selection, sign convention, scaling and the
 * disconnected-graph (multi-component) embedding.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define STREAM_NOISE 7u

/* find_components (src/components.rs:7-44): BFS in node order. */
int64_t or_find_components(int64_t n, const int64_t* indptr, const int32_t* indices,
                           int64_t* labels) {
    int64_t* queue = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) labels[i] = -1;
    int64_t nc = 0;
    for (int64_t s = 0; s < n; ++s) {
        if (labels[s] >= 0) continue;
        int64_t head = 0, tail = 0;
        labels[s] = nc;
        queue[tail++] = s;
        while (head < tail) {
            int64_t u = queue[head++];
            for (int64_t p = indptr[u]; p < indptr[u + 1]; ++p) {
                int64_t v = indices[p];
                if (labels[v] < 0) { labels[v] = nc; queue[tail++] = v; }
            }
        }
        ++nc;
    }
    free(queue);
    return nc;
}

/* select_eigenvectors (src/selection.rs:11-29): stable ascending sort of the
 * eigenvalues, skip position 0, take the next n_components columns. */
void or_select_eigenvectors(const double* evals, const double* evecs, int64_t n, int k,
                            int n_components, double* out) {
    int* idx = (int*)malloc(sizeof(int) * (size_t)k);
    for (int i = 0; i < k; ++i) idx[i] = i;
    for (int i = 1; i < k; ++i) {
        int t = idx[i], j = i - 1;
        while (j >= 0 && evals[idx[j]] > evals[t]) { idx[j + 1] = idx[j]; --j; }
        idx[j + 1] = t;
    }
    for (int64_t r = 0; r < n; ++r)
        for (int c = 0; c < n_components; ++c)
            out[(size_t)r * n_components + c] = evecs[(size_t)r * k + idx[1 + c]];
    free(idx);
}

/* normalize_signs (src/scaling.rs:62-74): the first element of largest
 * magnitude decides; a strictly larger |b| replaces the running candidate. */
void or_normalize_signs(double* coords, int64_t n, int d) {
    for (int c = 0; c < d; ++c) {
        if (n == 0) continue;
        double best = coords[c];
        for (int64_t i = 1; i < n; ++i) {
            double b = coords[(size_t)i * d + c];
            if (fabs(b) > fabs(best)) best = b;
        }
        if (best < 0.0 || (best == 0.0 && signbit(best)))
            for (int64_t i = 0; i < n; ++i) coords[(size_t)i * d + c] = -coords[(size_t)i * d + c];
    }
}

/* scale_coords (src/scaling.rs:9-21).  Returns -1 on an all-zero embedding. */
int or_scale_coords(const double* coords, int64_t n, int d, double max_coord, float* out) {
    double m = 0.0;
    size_t cnt = (size_t)n * d;
    for (size_t i = 0; i < cnt; ++i) if (fabs(coords[i]) > m) m = fabs(coords[i]);
    if (m <= 0.0) return -1;
    double ex = max_coord / m;
    for (size_t i = 0; i < cnt; ++i) out[i] = (float)(coords[i] * ex);
    return 0;
}

static void laplacian_of(int64_t n, const int64_t* ip, const int32_t* ix, const float* w,
                         int mode, int64_t** lip, int32_t** lix, double** ld, double** sq) {
    double* deg = (double*)malloc(sizeof(double) * (size_t)n);
    *sq = (double*)malloc(sizeof(double) * (size_t)n);
    or_compute_degrees(n, ip, ix, w, mode, deg, *sq);
    double* isd = (double*)malloc(sizeof(double) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) isd[i] = (*sq)[i] > 0.0 ? 1.0 / (*sq)[i] : 0.0;
    int64_t nnz = or_laplacian_nnz(n, ip, ix);
    *lip = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    *lix = (int32_t*)malloc(sizeof(int32_t) * (size_t)nnz);
    *ld = (double*)malloc(sizeof(double) * (size_t)nnz);
    or_build_laplacian(n, ip, ix, w, isd, *lip, *lix, *ld);
    free(deg); free(isd);
}

/* embed_disconnected (src/multi_component.rs:16-74) and helpers. */
static int embed_disconnected(int64_t n, const int64_t* ip, const int32_t* ix, const float* w,
                              const int64_t* labels, int64_t nc, int dims, uint64_t seed,
                              const float* raw, int64_t nf, int mode, int64_t dthr, double tol,
                              double* result) {
    /* members per component, ascending node order */
    int64_t* counts = (int64_t*)calloc((size_t)nc, sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) counts[labels[i]]++;
    int64_t* start = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nc + 1));
    start[0] = 0;
    for (int64_t c = 0; c < nc; ++c) start[c + 1] = start[c] + counts[c];
    int64_t* members = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int64_t* fill = (int64_t*)calloc((size_t)nc, sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) members[start[labels[i]] + fill[labels[i]]++] = i;
    double* meta = (double*)calloc((size_t)nc * dims, sizeof(double));
    int status = 0;
    if (nc > 2 * dims) {
        if (!raw) { status = 2; goto cleanup; }
        /* spectral_meta_embedding (multi_component.rs:143-220) */
        double* cen = (double*)calloc((size_t)nc * nf, sizeof(double));
        for (int64_t c = 0; c < nc; ++c) {
            for (int64_t q = start[c]; q < start[c + 1]; ++q)
                for (int64_t f = 0; f < nf; ++f) cen[c * nf + f] += (double)raw[members[q] * nf + f];
            double inv = 1.0 / (double)counts[c];
            for (int64_t f = 0; f < nf; ++f) cen[c * nf + f] *= inv;
        }
        double* aff = (double*)malloc(sizeof(double) * (size_t)nc * nc);
        for (int64_t i = 0; i < nc; ++i)
            for (int64_t j = 0; j < nc; ++j) {
                double d2 = 0.0;
                for (int64_t f = 0; f < nf; ++f) {
                    double d = cen[i * nf + f] - cen[j * nf + f];
                    d2 += d * d;
                }
                aff[i * nc + j] = exp(-d2);
            }
        double* deg = (double*)malloc(sizeof(double) * (size_t)nc);
        double* isd = (double*)malloc(sizeof(double) * (size_t)nc);
        double* sq = (double*)malloc(sizeof(double) * (size_t)nc);
        for (int64_t i = 0; i < nc; ++i) {
            double s = 0.0;
            for (int64_t j = 0; j < nc; ++j) if (j != i) s += aff[i * nc + j];
            deg[i] = s;
            isd[i] = s > 0.0 ? 1.0 / sqrt(s) : 0.0;
            sq[i] = sqrt(s);
        }
        int64_t* lip = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nc + 1));
        int32_t* lix = (int32_t*)malloc(sizeof(int32_t) * (size_t)nc * nc);
        double* ld = (double*)malloc(sizeof(double) * (size_t)nc * nc);
        int64_t o = 0;
        lip[0] = 0;
        for (int64_t i = 0; i < nc; ++i) {
            for (int64_t j = 0; j < nc; ++j) {
                lix[o] = (int32_t)j;
                ld[o] = (i == j) ? 1.0 : -isd[i] * aff[i * nc + j] * isd[j];
                ++o;
            }
            lip[i + 1] = o;
        }
        double* ev = (double*)malloc(sizeof(double) * (size_t)(dims + 1));
        double* evec = (double*)malloc(sizeof(double) * (size_t)nc * (dims + 1));
        or_solve_eigenproblem(nc, lip, lix, ld, dims, seed, sq, dthr, tol, ev, evec, NULL);
        or_select_eigenvectors(ev, evec, nc, dims + 1, dims, meta);
        double mx = 0.0;
        for (int64_t i = 0; i < nc * dims; ++i) if (fabs(meta[i]) > mx) mx = fabs(meta[i]);
        if (mx == 0.0) status = 2;
        else for (int64_t i = 0; i < nc * dims; ++i) meta[i] /= mx;
        /* the eigenvector signs are the eigensolver's choice (faer's in the
         * reference, unpinned): fixed here by the normalize_signs rule */
        or_normalize_signs(meta, nc, dims);
        free(cen); free(aff); free(deg); free(isd); free(sq); free(lip); free(lix); free(ld);
        free(ev); free(evec);
        if (status) goto cleanup;
    } else {
        /* orthogonal_meta_embedding (multi_component.rs:127-139) */
        for (int64_t r = 0; r < nc; ++r) meta[r * dims + r % dims] = r < dims ? 1.0 : -1.0;
    }
    /* compute_data_range (multi_component.rs:224-255) */
    double* range = (double*)malloc(sizeof(double) * (size_t)nc);
    for (int64_t c = 0; c < nc; ++c) {
        if (nc == 1) { range[c] = 1.0; continue; }
        double md = INFINITY;
        for (int64_t i = 0; i < nc; ++i) {
            if (i == c) continue;
            double s = 0.0;
            for (int d = 0; d < dims; ++d) {
                double df = meta[c * dims + d] - meta[i * dims + d];
                s += df * df;
            }
            s = sqrt(s);
            if (s > 0.0 && s < md) md = s;
        }
        range[c] = isfinite(md) ? md / 2.0 : 1.0;
    }
    /* embed_single_component (multi_component.rs:261-312) */
    int64_t* lookup = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) lookup[i] = -1;
    for (int64_t c = 0; c < nc && status == 0; ++c) {
        int64_t sz = counts[c];
        const int64_t* mem = members + start[c];
        if (sz < 2 * dims) {
            for (int64_t q = 0; q < sz; ++q)
                for (int d = 0; d < dims; ++d) result[mem[q] * dims + d] = meta[c * dims + d];
            continue;
        }
        for (int64_t q = 0; q < sz; ++q) lookup[mem[q]] = q;
        int64_t* sip = (int64_t*)malloc(sizeof(int64_t) * (size_t)(sz + 1));
        int64_t cap = 0;
        for (int64_t q = 0; q < sz; ++q) cap += ip[mem[q] + 1] - ip[mem[q]];
        int32_t* six = (int32_t*)malloc(sizeof(int32_t) * (size_t)(cap > 0 ? cap : 1));
        float* sw = (float*)malloc(sizeof(float) * (size_t)(cap > 0 ? cap : 1));
        int64_t o = 0;
        sip[0] = 0;
        for (int64_t q = 0; q < sz; ++q) {
            int64_t u = mem[q];
            for (int64_t p = ip[u]; p < ip[u + 1]; ++p)
                if (lookup[ix[p]] >= 0) { six[o] = (int32_t)lookup[ix[p]]; sw[o] = w[p]; ++o; }
            sip[q + 1] = o;
        }
        int64_t *lip; int32_t *lix; double *ld, *sq;
        laplacian_of(sz, sip, six, sw, mode, &lip, &lix, &ld, &sq);
        double* ev = (double*)malloc(sizeof(double) * (size_t)(dims + 1));
        double* evec = (double*)malloc(sizeof(double) * (size_t)sz * (dims + 1));
        double* co = (double*)malloc(sizeof(double) * (size_t)sz * dims);
        or_solve_eigenproblem(sz, lip, lix, ld, dims, seed, sq, dthr, tol, ev, evec, NULL);
        or_select_eigenvectors(ev, evec, sz, dims + 1, dims, co);
        or_normalize_signs(co, sz, dims);
        double mx = 0.0;
        for (int64_t i = 0; i < sz * dims; ++i) if (fabs(co[i]) > mx) mx = fabs(co[i]);
        if (mx > 0.0) for (int64_t i = 0; i < sz * dims; ++i) co[i] *= range[c] / mx;
        for (int64_t q = 0; q < sz; ++q)
            for (int d = 0; d < dims; ++d) result[mem[q] * dims + d] = co[q * dims + d] + meta[c * dims + d];
        for (int64_t q = 0; q < sz; ++q) lookup[mem[q]] = -1;
        free(sip); free(six); free(sw); free(lip); free(lix); free(ld); free(sq);
        free(ev); free(evec); free(co);
    }
    free(lookup); free(range);
cleanup:
    free(counts); free(start); free(members); free(fill); free(meta);
    return status;
}

/* spectral_init (src/lib.rs:198-258).  Error codes: 2 InvalidGraph,
 * 3 TooFewNodes.  add_noise = 0 returns the pre-noise scaled coordinates. */
int or_spectral_init(int64_t n, const int64_t* indptr, const int32_t* indices,
                     const float* data, int n_components, uint64_t seed,
                     const float* raw_data, int64_t n_features, int mode,
                     int64_t dense_threshold, double lobpcg_tol, int add_noise,
                     float* out) {
    if (n == 0 || n_components == 0) return 2;
    if (n <= n_components) return 3;
    int64_t* labels = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int64_t nc = or_find_components(n, indptr, indices, labels);
    double* coords = (double*)calloc((size_t)n * n_components, sizeof(double));
    int status = 0;
    if (nc > 1) {
        status = embed_disconnected(n, indptr, indices, data, labels, nc, n_components, seed,
                                    raw_data, n_features, mode, dense_threshold, lobpcg_tol,
                                    coords);
    } else {
        int64_t *lip; int32_t *lix; double *ld, *sq;
        laplacian_of(n, indptr, indices, data, mode, &lip, &lix, &ld, &sq);
        int k = n_components + 1;
        double* ev = (double*)malloc(sizeof(double) * (size_t)k);
        double* evec = (double*)malloc(sizeof(double) * (size_t)n * k);
        or_solve_eigenproblem(n, lip, lix, ld, n_components, seed, sq, dense_threshold,
                              lobpcg_tol, ev, evec, NULL);
        or_select_eigenvectors(ev, evec, n, k, n_components, coords);
        or_normalize_signs(coords, n, n_components);
        free(lip); free(lix); free(ld); free(sq); free(ev); free(evec);
    }
    if (status == 0) {
        if (or_scale_coords(coords, n, n_components, 10.0, out) != 0) status = 2;
        else if (add_noise)
            for (int64_t i = 0; i < n * n_components; ++i)
                out[i] += (float)(1e-4 * or_normal(seed, STREAM_NOISE, (uint64_t)i));
    }
    free(labels); free(coords);
    return status;
}
