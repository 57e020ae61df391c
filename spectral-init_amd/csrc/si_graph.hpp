// si_graph.hpp -- the fuzzy k-NN graph builder (k_graph.hip): device CSR
// graph and the launchers of its steps (reference fixture pipeline
// tests/generate_fixtures.py:70-182).
#pragma once

#include "si_host.hpp"

namespace si {

// CSR graph on the device: int64 row pointers, int32 columns ascending within
// each row, float32 weights.
struct DevGraph {
    int64_t n = 0;
    int64_t nnz = 0;
    DevBuf<int64_t> ip;
    DevBuf<int32_t> ix;
    DevBuf<float> w;
};

// Step 1: exact k-NN of n points [n x d] (precision 0 = float32, 1 = float64;
// device array), itself excluded; idx/dist [n x k] on the device.
void knn_exact_device(const void* points_dev, int precision, int64_t n, int64_t d, int64_t k,
                      int32_t* idx, float* dist, hipStream_t s);
// Testing seam of step 1: 0 automatic (float32 points with k <= 16 take the
// MFMA filter and the certified FP64 re-rank), 1 the direct FP64 kernel only,
// 2 the filter with every query re-run by the direct kernel; and the number
// of queries the last filtered call re-ran directly.
void knn_set_path(int mode);
int64_t knn_direct_count();
// Step 2: sigma, rho [n] from the distances [n x k] (device arrays).
void smooth_knn_device(const float* dist, int64_t n, int64_t k, float* sig, float* rho,
                       hipStream_t s);
// Step 3: membership values [n x k] (device arrays).
void membership_device(const int32_t* idx, const float* dist, const float* sig, const float* rho,
                       int64_t n, int64_t k, float* vals, hipStream_t s);
// Steps 4-5a: fuzzy union, zeros dropped, pruned below max/n_epochs when
// n_epochs > 0.
void fuzzy_union_device(const int32_t* idx, const float* vals, int64_t n, int64_t k,
                        int n_epochs, DevGraph& g, hipStream_t s);

}  // namespace si
