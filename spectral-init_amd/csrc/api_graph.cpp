// api_graph.cpp -- C ABI of the fuzzy k-NN graph builder
// (include/spectral_mi355x_graph.h): host arrays in and out, the steps on the
// device (k_graph.hip).
#include <memory>

#include "si_abi.hpp"
#include "si_graph.hpp"
#include "../../include/spectral_mi355x_graph.h"

struct si_graph {
    si::DevGraph g;
    int device = 0;
};

namespace {

template <typename F>
si_status gguard(si_context* ctx, F&& f) {
    if (!ctx) return SI_ERR_INVALID_ARGUMENT;
    try {
        SI_HIP(hipSetDevice(ctx->c.device));
        f();
        ctx->c.last_error.clear();
        return SI_OK;
    } catch (const si::Error& e) {
        ctx->c.last_error = e.what();
        return (si_status)e.code;
    } catch (const std::bad_alloc&) {
        ctx->c.last_error = "host allocation failed";
        return SI_ERR_INTERNAL;
    } catch (const std::exception& e) {
        ctx->c.last_error = e.what();
        return SI_ERR_INTERNAL;
    }
}

int epochs_rule(int32_t n_epochs, int64_t n) {
    // generate_fixtures.py:171-174: 500 for n <= 10000, else 200
    if (n_epochs < 0) return n <= 10000 ? 500 : 200;
    return n_epochs;
}

size_t point_bytes(int32_t precision) { return precision == SI_F64 ? sizeof(double) : sizeof(float); }

}  // namespace

extern "C" {

si_status si_knn_exact(si_context* ctx, const void* points, int32_t precision, int64_t n,
                       int64_t d, int64_t k, int32_t* knn_indices, float* knn_dists) {
    return gguard(ctx, [&] {
        SI_CHECK(points && knn_indices && knn_dists, si::kErrInvalidArgument,
                 "knn_exact: null argument");
        SI_CHECK(precision == SI_F32 || precision == SI_F64, si::kErrInvalidArgument,
                 "knn_exact: precision must be SI_F32 or SI_F64");
        SI_CHECK(n >= 2 && d >= 1 && k >= 1, si::kErrInvalidArgument,
                 "knn_exact: need n >= 2, d >= 1, k >= 1");
        hipStream_t s = ctx->c.stream;
        si::DevBuf<char> px((size_t)n * d * point_bytes(precision));
        si::DevBuf<int32_t> idx((size_t)n * k);
        si::DevBuf<float> dist((size_t)n * k);
        SI_HIP(hipMemcpyAsync(px.get(), points, (size_t)n * d * point_bytes(precision),
                              hipMemcpyHostToDevice, s));
        si::knn_exact_device(px.get(), precision == SI_F64 ? 1 : 0, n, d, k, idx, dist, s);
        SI_HIP(hipMemcpyAsync(knn_indices, idx.get(), sizeof(int32_t) * n * k,
                              hipMemcpyDeviceToHost, s));
        SI_HIP(hipMemcpyAsync(knn_dists, dist.get(), sizeof(float) * n * k, hipMemcpyDeviceToHost,
                              s));
        SI_HIP(hipStreamSynchronize(s));
    });
}

si_status si_smooth_knn_dist(si_context* ctx, const float* knn_dists, int64_t n, int64_t k,
                             float* sigmas, float* rhos) {
    return gguard(ctx, [&] {
        SI_CHECK(knn_dists && sigmas && rhos, si::kErrInvalidArgument,
                 "smooth_knn_dist: null argument");
        SI_CHECK(n >= 1 && k >= 2, si::kErrInvalidArgument, "smooth_knn_dist: need n >= 1, k >= 2");
        hipStream_t s = ctx->c.stream;
        si::DevBuf<float> dist((size_t)n * k), sig(n), rho(n);
        SI_HIP(hipMemcpyAsync(dist.get(), knn_dists, sizeof(float) * n * k, hipMemcpyHostToDevice,
                              s));
        si::smooth_knn_device(dist, n, k, sig, rho, s);
        SI_HIP(hipMemcpyAsync(sigmas, sig.get(), sizeof(float) * n, hipMemcpyDeviceToHost, s));
        SI_HIP(hipMemcpyAsync(rhos, rho.get(), sizeof(float) * n, hipMemcpyDeviceToHost, s));
        SI_HIP(hipStreamSynchronize(s));
    });
}

si_status si_membership_strengths(si_context* ctx, const int32_t* knn_indices,
                                  const float* knn_dists, const float* sigmas, const float* rhos,
                                  int64_t n, int64_t k, float* vals) {
    return gguard(ctx, [&] {
        SI_CHECK(knn_indices && knn_dists && sigmas && rhos && vals, si::kErrInvalidArgument,
                 "membership_strengths: null argument");
        SI_CHECK(n >= 1 && k >= 1, si::kErrInvalidArgument, "membership_strengths: need n, k >= 1");
        hipStream_t s = ctx->c.stream;
        si::DevBuf<int32_t> idx((size_t)n * k);
        si::DevBuf<float> dist((size_t)n * k), sig(n), rho(n), v((size_t)n * k);
        SI_HIP(hipMemcpyAsync(idx.get(), knn_indices, sizeof(int32_t) * n * k,
                              hipMemcpyHostToDevice, s));
        SI_HIP(hipMemcpyAsync(dist.get(), knn_dists, sizeof(float) * n * k, hipMemcpyHostToDevice,
                              s));
        SI_HIP(hipMemcpyAsync(sig.get(), sigmas, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SI_HIP(hipMemcpyAsync(rho.get(), rhos, sizeof(float) * n, hipMemcpyHostToDevice, s));
        si::membership_device(idx, dist, sig, rho, n, k, v, s);
        SI_HIP(hipMemcpyAsync(vals, v.get(), sizeof(float) * n * k, hipMemcpyDeviceToHost, s));
        SI_HIP(hipStreamSynchronize(s));
    });
}

si_status si_fuzzy_union(si_context* ctx, const int32_t* knn_indices, const float* vals,
                         int64_t n, int64_t k, int32_t n_epochs, si_graph** out) {
    if (out) *out = nullptr;
    return gguard(ctx, [&] {
        SI_CHECK(knn_indices && vals && out, si::kErrInvalidArgument, "fuzzy_union: null argument");
        SI_CHECK(n >= 1 && k >= 1, si::kErrInvalidArgument, "fuzzy_union: need n, k >= 1");
        hipStream_t s = ctx->c.stream;
        si::DevBuf<int32_t> idx((size_t)n * k);
        si::DevBuf<float> v((size_t)n * k);
        SI_HIP(hipMemcpyAsync(idx.get(), knn_indices, sizeof(int32_t) * n * k,
                              hipMemcpyHostToDevice, s));
        SI_HIP(hipMemcpyAsync(v.get(), vals, sizeof(float) * n * k, hipMemcpyHostToDevice, s));
        auto g = std::make_unique<si_graph>();
        g->device = ctx->c.device;
        si::fuzzy_union_device(idx, v, n, k, epochs_rule(n_epochs, n), g->g, s);
        *out = g.release();
    });
}

si_status si_fuzzy_graph_build(si_context* ctx, const void* points, int32_t precision, int64_t n,
                               int64_t d, int64_t k, int32_t n_epochs, si_graph** out) {
    if (out) *out = nullptr;
    return gguard(ctx, [&] {
        SI_CHECK(points && out, si::kErrInvalidArgument, "fuzzy_graph_build: null argument");
        SI_CHECK(precision == SI_F32 || precision == SI_F64, si::kErrInvalidArgument,
                 "fuzzy_graph_build: precision must be SI_F32 or SI_F64");
        SI_CHECK(n >= 2 && d >= 1 && k >= 2, si::kErrInvalidArgument,
                 "fuzzy_graph_build: need n >= 2, d >= 1, k >= 2");
        hipStream_t s = ctx->c.stream;
        si::DevBuf<char> px((size_t)n * d * point_bytes(precision));
        SI_HIP(hipMemcpyAsync(px.get(), points, (size_t)n * d * point_bytes(precision),
                              hipMemcpyHostToDevice, s));
        si::DevBuf<int32_t> idx((size_t)n * k);
        si::DevBuf<float> dist((size_t)n * k), sig(n), rho(n), v((size_t)n * k);
        si::knn_exact_device(px.get(), precision == SI_F64 ? 1 : 0, n, d, k, idx, dist, s);
        px.release();
        si::smooth_knn_device(dist, n, k, sig, rho, s);
        si::membership_device(idx, dist, sig, rho, n, k, v, s);
        auto g = std::make_unique<si_graph>();
        g->device = ctx->c.device;
        si::fuzzy_union_device(idx, v, n, k, epochs_rule(n_epochs, n), g->g, s);
        *out = g.release();
    });
}

si_status si_graph_info(const si_graph* g, int64_t* n, int64_t* nnz) {
    if (!g) return SI_ERR_INVALID_ARGUMENT;
    if (n) *n = g->g.n;
    if (nnz) *nnz = g->g.nnz;
    return SI_OK;
}

si_status si_graph_copy(si_context* ctx, const si_graph* g, int64_t* indptr, int32_t* indices,
                        float* data) {
    return gguard(ctx, [&] {
        SI_CHECK(g && indptr, si::kErrInvalidArgument, "graph_copy: null argument");
        SI_CHECK(g->g.nnz == 0 || (indices && data), si::kErrInvalidArgument,
                 "graph_copy: null argument");
        hipStream_t s = ctx->c.stream;
        SI_HIP(hipMemcpyAsync(indptr, g->g.ip.get(), sizeof(int64_t) * (g->g.n + 1),
                              hipMemcpyDeviceToHost, s));
        if (g->g.nnz > 0) {
            SI_HIP(hipMemcpyAsync(indices, g->g.ix.get(), sizeof(int32_t) * g->g.nnz,
                                  hipMemcpyDeviceToHost, s));
            SI_HIP(hipMemcpyAsync(data, g->g.w.get(), sizeof(float) * g->g.nnz,
                                  hipMemcpyDeviceToHost, s));
        }
        SI_HIP(hipStreamSynchronize(s));
    });
}

void si_graph_destroy(si_graph* g) {
    if (!g) return;
    int cur = 0;
    if (hipGetDevice(&cur) == hipSuccess && cur != g->device) {
        (void)hipSetDevice(g->device);
        delete g;
        (void)hipSetDevice(cur);
        return;
    }
    delete g;
}

}  // extern "C"
