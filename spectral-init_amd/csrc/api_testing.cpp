// api_testing.cpp -- kernel-level seams (include/spectral_mi355x_testing.h),
// the counterpart of the reference's `testing`-feature re-exports
// (src/lib.rs:27-162, src/operator.rs:16-32): they let the parity tests drive
// single device kernels against the CPU oracle.
#include <cmath>
#include <cstring>
#include <vector>

#include "si_abi.hpp"
#include "si_graph.hpp"
#include "../../include/spectral_mi355x_testing.h"

namespace {
template <typename F>
si_status tguard(si_context* ctx, F&& f) {
    try {
        if (ctx) SI_HIP(hipSetDevice(ctx->c.device));
        f();
        return SI_OK;
    } catch (const si::Error& e) {
        if (ctx) ctx->c.last_error = e.what();
        return (si_status)e.code;
    } catch (const std::exception& e) {
        if (ctx) ctx->c.last_error = e.what();
        return SI_ERR_INTERNAL;
    }
}
}  // namespace

namespace si {
extern bool g_batch_graphs;
extern bool g_device_output;
extern BatchStats g_last_batch;
}  // namespace si

extern "C" {

si_status si_test_gram(si_context* ctx, int32_t precision, int64_t n, const void* a, int32_t lda,
                       int32_t p, const void* b, int32_t ldb, int32_t q, double* g_host) {
    if (!ctx || !a || !b || !g_host) return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] {
        hipStream_t s = ctx->c.stream;
        si::DevBuf<double> g((size_t)p * q), ws(si::gram_workspace_doubles(n, p, q));
        if (precision == SI_F32)
            si::gram_f32(n, (const float*)a, lda, p, (const float*)b, ldb, q, g, ws, s);
        else
            si::gram_f64(n, (const double*)a, lda, p, (const double*)b, ldb, q, g, ws, s);
        si::to_host(g_host, g.get(), (size_t)p * q, s);
    });
}

si_status si_test_block_mul(si_context* ctx, int32_t precision, int64_t n, const void* a,
                            int32_t lda, int32_t p, const double* c_host, int32_t q, void* out,
                            int32_t ldo) {
    if (!ctx || !a || !c_host || !out) return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] {
        hipStream_t s = ctx->c.stream;
        si::DevBuf<double> c((size_t)p * q);
        si::to_device(c.get(), c_host, (size_t)p * q, s);
        if (precision == SI_F32)
            si::block_mul_f32(n, (const float*)a, lda, p, c, q, nullptr, 0, 0, nullptr, (float*)out, ldo, s);
        else
            si::block_mul_f64(n, (const double*)a, lda, p, c, q, nullptr, 0, 0, nullptr, (double*)out, ldo, s);
        SI_HIP(hipStreamSynchronize(s));
    });
}

si_status si_test_small_sym_eig(si_context* ctx, int32_t n, const double* a_host, double* w_host,
                                double* v_host) {
    if (!ctx || !a_host || !w_host || !v_host) return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] {
        hipStream_t s = ctx->c.stream;
        si::DevBuf<double> a((size_t)n * n), w(n), v((size_t)n * n);
        si::to_device(a.get(), a_host, (size_t)n * n, s);
        si::small_sym_eig(n, a, n, w, v, s);
        si::to_host(w_host, w.get(), n, s);
        si::to_host(v_host, v.get(), (size_t)n * n, s);
    });
}

si_status si_test_small_gen_eig(si_context* ctx, int32_t n, const double* a_host,
                                const double* b_host, double* w_host, double* c_host) {
    if (!ctx || !a_host || !b_host || !w_host || !c_host) return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] {
        hipStream_t s = ctx->c.stream;
        si::DevBuf<double> a((size_t)n * n), b((size_t)n * n), w(n), c((size_t)n * n);
        si::DevBuf<int> st(1);
        si::to_device(a.get(), a_host, (size_t)n * n, s);
        si::to_device(b.get(), b_host, (size_t)n * n, s);
        si::small_gen_eig(n, a, b, w, c, st, s);
        int h = 0;
        si::to_host(&h, st.get(), 1, s);
        SI_CHECK(h == 0, si::kErrInternal, "small_gen_eig: B is not positive definite");
        si::to_host(w_host, w.get(), n, s);
        si::to_host(c_host, c.get(), (size_t)n * n, s);
    });
}

si_status si_test_fill_normal(si_context* ctx, double* x, int64_t n, int32_t k, uint64_t seed,
                              uint32_t stream) {
    if (!ctx || !x) return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] {
        si::fill_normal_f64(x, n, k, k, seed, stream, ctx->c.stream);
        SI_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

si_status si_test_fill_normal_f32(si_context* ctx, float* x, int64_t n, int32_t k, uint64_t seed,
                                  uint32_t stream) {
    if (!ctx || !x || n < 0 || k < 1) return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] {
        si::fill_normal_f32(x, n, k, k, seed, stream, ctx->c.stream);
        SI_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

si_status si_test_staged_roundtrip(si_context* ctx, const void* in, void* out, int64_t bytes) {
    if (!ctx || !in || !out || bytes < 1) return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] {
        si::StreamBuf<char> d((size_t)bytes, ctx->c.stream);
        si::upload_to_device(d.get(), in, (size_t)bytes, ctx->c.stream);
        si::download_to_host(out, d.get(), (size_t)bytes, ctx->c.stream);
    });
}

si_status si_test_fast_solve(si_context* ctx, const si_laplacian* lap, int32_t n_components,
                             uint64_t seed, const si_config* cfg, double* evals, double* evecs,
                             si_report* rep) {
    if (!ctx || !lap || !evals || !evecs) return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] {
        si_config d;
        si_config_default(&d);
        if (!cfg) cfg = &d;
        si::FastOptions o;
        o.block_size = cfg->block_size;
        o.cheb_degree = cfg->cheb_degree > 0 ? cfg->cheb_degree : 0;
        o.max_outer = cfg->max_outer > 0 ? cfg->max_outer : 60;
        o.tol = cfg->tol > 0 ? cfg->tol : 1e-6;
        o.warm_start = cfg->warm_start;
        si::SolveStats st;
        const int k = n_components + 1;
        hipStream_t s = ctx->c.stream;
        si::DevLaplacian& ml = const_cast<si::DevLaplacian&>(lap->L);
        const int tw = si::fast_tune_widths(ml.n, k, o.block_size) & ~ml.tuned;
        if (tw && ml.nnz >= (int64_t)1 << 20) si::spmm_autotune(ml, s, tw);
        si::DevBuf<double> vi((size_t)lap->L.n * k), wsi;
        if (o.warm_start) {
            wsi.alloc((size_t)lap->L.n * k);
            si::rows_to_internal(lap->L, o.warm_start, k, wsi.get(), k, k, s);
            o.warm_start = wsi.get();
        }
        si::EventTimer t(s);
        const bool ok = si::fast_solve(lap->L, n_components, seed, o, evals, vi, st, s);
        if (ok) si::rows_to_external(lap->L, vi.get(), k, evecs, k, k, s);
        const double ms = t.stop_ms();
        if (rep) {
            rep->level = 1;
            rep->restarts = 0;
            rep->block_size = st.block_size;
            rep->iterations = st.iterations;
            rep->spmm_calls = st.spmm_calls;
            rep->max_residual = st.max_residual;
            for (int i = 0; i < 8; ++i) rep->eigenvalues[i] = i <= n_components ? evals[i] : 0.0;
            rep->solve_ms = ms;
        }
        SI_CHECK(ok, si::kErrConvergence, "fast_solve produced no result");
    });
}

si_status si_test_band_solve(si_context* ctx, const si_laplacian* lap, double shift,
                             const double* x, double* y, int32_t k, int32_t* bandwidth) {
    if (!ctx || !lap || !x || !y || k < 1) return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] {
        hipStream_t s = ctx->c.stream;
        const int64_t n = lap->L.n;
        si::BandFactor F;
        const bool ok = si::band_factor(lap->L, shift, F, s);
        if (bandwidth) *bandwidth = ok ? F.b : -1;
        SI_CHECK(ok, si::kErrUnsupported, "band_solve: band too large or not positive definite");
        si::DevBuf<double> xi((size_t)n * k), yi((size_t)n * k);
        si::rows_to_internal(lap->L, x, k, xi.get(), k, k, s);
        si::band_solve(F, xi.get(), k, yi.get(), k, k, 1.0, s);
        si::rows_to_external(lap->L, yi.get(), k, y, k, k, s);
        SI_HIP(hipStreamSynchronize(s));
    });
}

si_status si_profile_spmm_begin(si_context* ctx) {
    if (!ctx) return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] {
        std::vector<hipEvent_t> pool = std::move(ctx->c.profile.pool);  // kept across regions
        ctx->c.profile = si::SpmmProfile{};
        ctx->c.profile.pool = std::move(pool);
        // create the events of a long region up front, outside the timed work
        while (ctx->c.profile.pool.size() < 1024) {
            hipEvent_t e = nullptr;
            SI_HIP(hipEventCreate(&e));
            ctx->c.profile.pool.push_back(e);
        }
        ctx->c.profile.enabled = true;
        si::spmm_profile_attach(&ctx->c.profile);
    });
}

si_status si_profile_spmm_end(si_context* ctx, si_spmm_profile* out) {
    if (!ctx || !out) return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] {
        si::spmm_profile_collect(&ctx->c.profile);
        si::spmm_profile_attach(nullptr);
        const si::SpmmProfile& p = ctx->c.profile;
        out->launches = p.launches;
        out->bytes = p.bytes;
        out->ms = p.ms;
        out->launches_f32 = p.launches_f32;
        out->bytes_f32 = p.bytes_f32;
        out->ms_f32 = p.ms_f32;
        ctx->c.profile.enabled = false;
    });
}

si_status si_profile_gather_probe(si_context* ctx, const si_laplacian* lap, int32_t mode,
                                  int32_t reps, double* us) {
    if (!ctx || !lap || !us) return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] { *us = si::gather_probe_us(lap->L, mode, reps, ctx->c.stream); });
}

si_status si_test_batch_plan(si_context* ctx, int32_t mode, int32_t update, int32_t nseg,
                             int32_t k, int32_t m, int32_t cap, int32_t steps,
                             int32_t max_rounds, double tol, double* a, double* best,
                             double* worst, int32_t* stall, int32_t* done, double* thk,
                             const int32_t* status, const double* theta, const double* norms2,
                             int32_t* ctl, int32_t* deg, double* coef) {
    if (!ctx || !a || !best || !worst || !stall || !done || !status || !theta || !norms2 || !ctl ||
        !deg || !coef || (mode == 1 && !thk))
        return SI_ERR_INVALID_ARGUMENT;
    if ((mode != 0 && mode != 1) || update < 0 || update > 2 || nseg < 1 || k < 1 || m < k ||
        m > 64 || steps < 2 || cap < 1)
        return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] {
        hipStream_t s = ctx->c.stream;
        const size_t ns = (size_t)nseg;
        si::DevBuf<double> dd(ns * 3 + ns * m + ns * m + ns * 2 * k);
        si::DevBuf<int32_t> di(ns * 4 + 4);
        const size_t ncoef = (size_t)steps * (ns + 1) * 3;
        si::DevBuf<double> dcoef(ncoef);
        si::BatchPlan p;
        p.nseg = nseg; p.k = k; p.m = m; p.cap = cap; p.steps = steps;
        p.max_rounds = max_rounds; p.tol = tol; p.b = 2.0;
        p.a = dd.get(); p.best = p.a + ns; p.worst = p.best + ns;
        p.thk = mode == 1 ? p.worst + ns : nullptr;
        double* th = p.worst + ns + ns * m;
        double* nr = th + ns * m;
        p.stall = di.get(); p.done = p.stall + ns; p.deg = p.done + ns; p.ctl = p.deg + ns;
        int32_t* st = p.ctl + 4;
        p.status = st; p.theta = th; p.norms2 = nr;
        si::to_device(p.a, a, ns, s);
        si::to_device(p.best, best, ns, s);
        si::to_device(p.worst, worst, ns, s);
        if (p.thk) si::to_device(p.thk, thk, ns * m, s);
        si::to_device(th, theta, ns * m, s);
        si::to_device(nr, norms2, ns * 2 * k, s);
        si::to_device(p.stall, stall, ns, s);
        si::to_device(p.done, done, ns, s);
        si::to_device(p.ctl, ctl, 4, s);
        si::to_device(st, status, ns, s);
        if (mode == 0) {
            float* c32 = reinterpret_cast<float*>(dcoef.get());
            si::batch_plan_f32(p, update, c32, s);
            std::vector<float> h(ncoef);
            si::to_host(h.data(), c32, ncoef, s);
            for (size_t i = 0; i < ncoef; ++i) coef[i] = h[i];
        } else {
            si::batch_plan_f64(p, update, dcoef.get(), s);
            si::to_host(coef, dcoef.get(), ncoef, s);
        }
        si::to_host(a, p.a, ns, s);
        si::to_host(best, p.best, ns, s);
        si::to_host(worst, p.worst, ns, s);
        if (p.thk) si::to_host(thk, p.thk, ns * m, s);
        si::to_host(stall, p.stall, ns, s);
        si::to_host(done, p.done, ns, s);
        si::to_host(deg, p.deg, ns, s);
        si::to_host(ctl, p.ctl, 4, s);
    });
}

si_status si_test_sinv_dense(int32_t enable) {
    si::sinv_force_dense(enable != 0);
    return SI_OK;
}

si_status si_test_batch_graphs(int32_t enable, int64_t* captured, int64_t* launches) {
    if (captured) *captured = si::g_last_batch.graphs_captured;
    if (launches) *launches = si::g_last_batch.graph_launches;
    if (enable >= 0) si::g_batch_graphs = enable != 0;
    return SI_OK;
}

si_status si_test_spmm_shapes(si_context* ctx, si_laplacian* lap, int32_t shape16,
                              int32_t shape32, int32_t autotune, int32_t* out16, int32_t* out32) {
    if (!ctx || !lap) return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] {
        const int nc = si::spmm_shape_count();
        if (autotune) {
            si::spmm_autotune(lap->L, ctx->c.stream);
        } else {
            SI_CHECK(shape16 >= 0 && shape16 < nc && shape32 >= 0 && shape32 < nc,
                     si::kErrInvalidArgument, "spmm shape out of range");
            lap->L.shape16 = shape16;
            lap->L.shape32 = shape32;
            lap->L.tuned = si::kTuned16 | si::kTuned32;
        }
        if (out16) *out16 = lap->L.shape16;
        if (out32) *out32 = lap->L.shape32;
    });
}

si_status si_test_rcm_order(int64_t n, const int64_t* indptr, const int32_t* indices,
                            int32_t threads, int32_t* perm) {
    if (n < 1 || !indptr || !perm || (indptr[n] > 0 && !indices)) return SI_ERR_INVALID_ARGUMENT;
    try {
        const std::vector<int32_t> p = si::rcm_order(n, indptr, indices, threads);
        std::copy(p.begin(), p.end(), perm);
        return SI_OK;
    } catch (...) {
        return SI_ERR_INTERNAL;
    }
}

namespace {
struct DevPattern {
    si::DevBuf<int64_t> ip;
    si::DevBuf<int32_t> ix;
    DevPattern(int64_t n, const int64_t* indptr, const int32_t* indices, hipStream_t s)
        : ip(n + 1), ix(indptr[n]) {
        si::to_device(ip.get(), indptr, n + 1, s);
        if (indptr[n] > 0) si::to_device(ix.get(), indices, indptr[n], s);
    }
};
}  // namespace

si_status si_test_components_device(si_context* ctx, int64_t n, const int64_t* indptr,
                                    const int32_t* indices, int32_t* labels, int64_t* n_comp) {
    if (!ctx || n < 1 || !indptr || !labels || !n_comp || (indptr[n] > 0 && !indices))
        return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] {
        hipStream_t s = ctx->c.stream;
        DevPattern g(n, indptr, indices, s);
        si::DevBuf<int32_t> lab(n);
        *n_comp = si::components_device(n, g.ip.get(), g.ix.get(), lab.get(), nullptr, s);
        si::to_host(labels, lab.get(), n, s);
    });
}

si_status si_test_rcm_order_device(si_context* ctx, int64_t n, const int64_t* indptr,
                                   const int32_t* indices, int32_t* perm) {
    if (!ctx || n < 1 || !indptr || !perm || (indptr[n] > 0 && !indices))
        return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] {
        hipStream_t s = ctx->c.stream;
        DevPattern g(n, indptr, indices, s);
        si::DevBuf<int32_t> p(n), ip(n);
        si::rcm_order_device(n, g.ip.get(), g.ix.get(), p.get(), ip.get(), s);
        si::to_host(perm, p.get(), n, s);
    });
}

si_status si_test_laplacian_storage(si_context* ctx, const si_laplacian* lap, int32_t* rowptr,
                                    int32_t* col, float* a32, int32_t* perm, int32_t* order) {
    if (!ctx || !lap) return SI_ERR_INVALID_ARGUMENT;
    return tguard(ctx, [&] {
        hipStream_t s = ctx->c.stream;
        const si::DevLaplacian& L = lap->L;
        if (rowptr) si::to_host(rowptr, L.rowptr, L.n + 1, s);
        if (col && L.nnz) si::to_host(col, L.col, L.nnz, s);
        if (a32 && L.nnz) si::to_host(a32, L.a32, L.nnz, s);
        if (perm) {
            if (L.perm) si::to_host(perm, L.perm, L.n, s);
            else for (int64_t i = 0; i < L.n; ++i) perm[i] = (int32_t)i;
        }
        if (order) *order = L.order;
    });
}

si_status si_test_output_stage(int32_t device) {
    if (device != 0 && device != 1) return SI_ERR_INVALID_ARGUMENT;
    si::g_device_output = device != 0;
    return SI_OK;
}

si_status si_test_knn_path(int32_t mode, int64_t* last_direct) {
    if (mode < -1 || mode > 3) return SI_ERR_INVALID_ARGUMENT;
    if (mode >= 0) si::knn_set_path(mode);
    if (last_direct) *last_direct = si::knn_direct_count();
    return SI_OK;
}

}  // extern "C"
