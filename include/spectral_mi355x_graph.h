/*
 * spectral_mi355x_graph.h -- the fuzzy k-NN graph W that spectral_init takes as
 * input, built on the device.
 *
 * The reference takes W from its caller (umap-rs; src/lib.rs:198-204) and builds
 * its fixture graphs with umap-learn 0.5.11 in tests/generate_fixtures.py.  These
 * entry points are that pipeline, one per step, plus the whole chain:
 *   step 1  exact k-NN, self excluded            generate_fixtures.py:70-93
 *   step 2  smooth_knn_dist -> sigma, rho        generate_fixtures.py:96-111
 *   step 3  compute_membership_strengths         generate_fixtures.py:114-141
 *   step 4  fuzzy union A + A^T - A o A^T        generate_fixtures.py:144-161
 *   step 5a prune below max(W) / n_epochs        generate_fixtures.py:164-182
 *
 * All arrays here are host arrays, row-major; the chain keeps its
 * intermediates on the device.  Points are FP32 or FP64 (si_precision) and
 * distances are computed in FP64 either way; d <= 64, k <= 64.
 */
#ifndef SPECTRAL_MI355X_GRAPH_H
#define SPECTRAL_MI355X_GRAPH_H

#include "spectral_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A CSR graph held on the device (int64 row pointers, int32 columns sorted
 * within each row, float32 weights). */
typedef struct si_graph si_graph;

/* Step 1 (generate_fixtures.py:70-93): the k nearest neighbours of every point
 * by exact Euclidean distance, itself excluded, ascending; ties go to the
 * smaller index.  points [n x d], precision SI_F32 or SI_F64; outputs
 * knn_indices [n x k] (int32) and knn_dists [n x k] (float32). */
si_status si_knn_exact(si_context* ctx, const void* points, int32_t precision, int64_t n,
                       int64_t d, int64_t k, int32_t* knn_indices, float* knn_dists);

/* Step 2 (generate_fixtures.py:96-111; umap smooth_knn_dist with n_iter 64,
 * local_connectivity 1, bandwidth 1): rho_i = the first non-zero distance of
 * row i; sigma_i by bisection on sum_{j>=1} exp(-max(d_ij - rho_i, 0)/sigma)
 * = log2(k), floored at 1e-3 times the mean distance.  Like the fixtures,
 * a row whose first sum is at or below log2(k) gets sigma = inf. */
si_status si_smooth_knn_dist(si_context* ctx, const float* knn_dists, int64_t n, int64_t k,
                             float* sigmas, float* rhos);

/* Step 3 (generate_fixtures.py:114-141; umap compute_membership_strengths):
 * vals [n x k] = exp(-(d_ij - rho_i) / sigma_i), 1 where d_ij <= rho_i or
 * sigma_i = 0, 0 for a self or missing (-1) neighbour. */
si_status si_membership_strengths(si_context* ctx, const int32_t* knn_indices,
                                  const float* knn_dists, const float* sigmas, const float* rhos,
                                  int64_t n, int64_t k, float* vals);

/* Steps 4 and 5a (generate_fixtures.py:144-182): W = A + A^T - A o A^T of the
 * directed graph A (row i: (knn_indices[i, j], vals[i, j])), float32, zeros
 * dropped; then, for n_epochs > 0, entries below max(W) / n_epochs are
 * dropped (n_epochs < 0: the fixtures' rule, 500 for n <= 10000, else 200;
 * 0: no pruning). */
si_status si_fuzzy_union(si_context* ctx, const int32_t* knn_indices, const float* vals,
                         int64_t n, int64_t k, int32_t n_epochs, si_graph** out);

/* Steps 1-5a in one call, intermediates on the device. */
si_status si_fuzzy_graph_build(si_context* ctx, const void* points, int32_t precision, int64_t n,
                               int64_t d, int64_t k, int32_t n_epochs, si_graph** out);

si_status si_graph_info(const si_graph* g, int64_t* n, int64_t* nnz);
/* host copies: indptr [n + 1] (int64), indices [nnz] (int32), data [nnz] */
si_status si_graph_copy(si_context* ctx, const si_graph* g, int64_t* indptr, int32_t* indices,
                        float* data);
void si_graph_destroy(si_graph* g);

#ifdef __cplusplus
}
#endif
#endif
