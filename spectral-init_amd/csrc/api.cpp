// api.cpp -- the C ABI (include/spectral_mi355x.h) over the device solvers,
// plus the spectral_init pipeline around them (src/lib.rs:198-258).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <exception>
#include <memory>
#include <numeric>
#include <thread>

#include "si_abi.hpp"

namespace {

template <typename F>
si_status guard(si_context* ctx, F&& f) {
    try {
        if (ctx) SI_HIP(hipSetDevice(ctx->c.device));
        f();
        if (ctx) ctx->c.last_error.clear();
        return SI_OK;
    } catch (const si::Error& e) {
        if (ctx) ctx->c.last_error = e.what();
        return (si_status)e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->c.last_error = "host allocation failed";
        return SI_ERR_INTERNAL;
    } catch (const std::exception& e) {
        if (ctx) ctx->c.last_error = e.what();
        return SI_ERR_INTERNAL;
    }
}

void fill_report(si_report* rep, const si::SolveStats& st, const double* ev, int k, double ms) {
    if (!rep) return;
    rep->level = st.level;
    rep->restarts = st.restarts;
    rep->block_size = st.block_size;
    rep->iterations = st.iterations;
    rep->spmm_calls = st.spmm_calls;
    rep->max_residual = st.max_residual;
    for (int i = 0; i < 8; ++i) rep->eigenvalues[i] = i < k ? ev[i] : 0.0;
    rep->solve_ms = ms;
    if (rep->struct_size >= offsetof(si_report, level_ms) + sizeof(rep->level_ms)) {
        rep->passes = st.passes;
        rep->reserved = 0;
        for (int l = 0; l < 6; ++l) rep->level_ms[l] = st.level_ms[l];
    }
}

si::ChainOptions chain_options(const si_config* cfg) {
    si::ChainOptions o;
    si_config d;
    si_config_default(&d);
    if (!cfg) cfg = &d;
    o.solver = cfg->solver;
    o.tol = cfg->tol > 0 ? cfg->tol : 1e-5;
    o.dense_threshold = cfg->dense_threshold >= 0 ? cfg->dense_threshold : 2000;
    o.fast.block_size = cfg->block_size;
    o.fast.cheb_degree = cfg->cheb_degree > 0 ? cfg->cheb_degree : 0;
    o.fast.max_outer = cfg->max_outer > 0 ? cfg->max_outer : 60;
    o.fast.warm_start = cfg->warm_start;
    o.fast.tol = o.tol;
    return o;
}

// ---------------------------------------------------------------- host helpers
// find_components (src/components.rs:7-44): BFS labelling in node order.
int64_t find_components(int64_t n, const int64_t* ip, const int32_t* ix, std::vector<int64_t>& lab) {
    lab.assign(n, -1);
    int64_t nc = 0;
    std::vector<int64_t> q;
    q.reserve(n);
    for (int64_t s0 = 0; s0 < n; ++s0) {
        if (lab[s0] >= 0) continue;
        q.clear();
        lab[s0] = nc;
        q.push_back(s0);
        for (size_t h = 0; h < q.size(); ++h) {
            const int64_t u = q[h];
            for (int64_t p = ip[u]; p < ip[u + 1]; ++p) {
                const int64_t v = ix[p];
                if (lab[v] < 0) { lab[v] = nc; q.push_back(v); }
            }
        }
        ++nc;
    }
    return nc;
}

// select_eigenvectors (src/selection.rs:11-29) + normalize_signs
// (src/scaling.rs:62-74) on host copies of the eigenvectors.
void select_and_sign(const std::vector<double>& ev, const std::vector<double>& vec, int64_t n,
                     int k, int d, std::vector<double>& out) {
    SI_CHECK((int)ev.size() == k && k > d && vec.size() == (size_t)n * k, si::kErrInternal,
             "select_and_sign: eigenpairs missing");
    std::vector<int> idx(k);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return ev[a] < ev[b]; });
    out.assign((size_t)n * d, 0.0);
    for (int64_t r = 0; r < n; ++r)
        for (int c = 0; c < d; ++c) out[(size_t)r * d + c] = vec[(size_t)r * k + idx[1 + c]];
    for (int c = 0; c < d; ++c) {
        double best = out[c];
        for (int64_t r = 1; r < n; ++r) {
            const double b = out[(size_t)r * d + c];
            if (std::fabs(b) > std::fabs(best)) best = b;
        }
        if (std::signbit(best))
            for (int64_t r = 0; r < n; ++r) out[(size_t)r * d + c] = -out[(size_t)r * d + c];
    }
}

// The centroid of every component's raw data rows (spectral_meta_embedding,
// src/multi_component.rs:143-165): FP64 sums in member order, one component
// per task, on host threads that run beside the caller until join().  The
// caller keeps raw and members alive and unchanged until then.
struct CentroidJob {
    std::vector<double> cen;  // nc x n_features
    std::vector<std::thread> pool;
    std::atomic<int64_t> next{0};

    void start(const float* raw, int64_t nf, const std::vector<std::vector<int64_t>>& members) {
        const int64_t nc = (int64_t)members.size();
        cen.assign((size_t)nc * nf, 0.0);
        auto centroid = [this, raw, nf, &members](int64_t c) {
            double* cc = cen.data() + c * nf;
            for (int64_t node : members[c])
                for (int64_t f = 0; f < nf; ++f) cc[f] += (double)raw[node * nf + f];
            const double inv = 1.0 / (double)members[c].size();
            for (int64_t f = 0; f < nf; ++f) cc[f] *= inv;
        };
        // the caller's thread keeps a core for the device work it drives
        const unsigned hw = std::max(2u, std::thread::hardware_concurrency());
        const int nth = (int)std::min<int64_t>(nc, std::min(8u, hw - 1));
        for (int t = 0; t < nth; ++t)
            pool.emplace_back([this, nc, centroid] {
                for (int64_t c = next++; c < nc; c = next++) centroid(c);
            });
    }
    void join() {
        for (auto& th : pool)
            if (th.joinable()) th.join();
        pool.clear();
    }
    ~CentroidJob() { join(); }
};

// the caller's CSR on the device for one build (stream-ordered scratch)
// members[c] = the nodes of component c in increasing order (a counting sort
// by label; on host threads, each over a contiguous range of nodes, when the
// per-thread counts stay small).
template <typename Label>
void member_lists(const std::vector<Label>& labels, int64_t nc,
                  std::vector<std::vector<int64_t>>& members) {
    const int64_t n = (int64_t)labels.size();
    members.assign(nc, {});
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nth = (int)std::min<int64_t>({8, (int64_t)hw, std::max<int64_t>(1, n / 65536)});
    if (nth <= 1 || nc > 65536) {
        std::vector<int64_t> cnt(nc, 0);
        for (int64_t i = 0; i < n; ++i) ++cnt[labels[i]];
        for (int64_t c = 0; c < nc; ++c) members[c].reserve(cnt[c]);
        for (int64_t i = 0; i < n; ++i) members[labels[i]].push_back(i);
        return;
    }
    // thread t takes nodes [t n / nth, (t + 1) n / nth); its nodes of component
    // c go after those of threads 0..t-1, so every list stays increasing
    std::vector<int64_t> pos((size_t)nth * nc, 0);
    auto on_threads = [&](auto&& body) {
        std::vector<std::thread> pool;
        try {
            for (int u = 0; u < nth; ++u)
                pool.emplace_back([&, u] { body(u, n * u / nth, n * (u + 1) / nth); });
        } catch (...) {
            for (auto& th : pool) th.join();
            throw;
        }
        for (auto& th : pool) th.join();
    };
    on_threads([&](int u, int64_t i0, int64_t i1) {
        int64_t* cu = pos.data() + (size_t)u * nc;
        for (int64_t i = i0; i < i1; ++i) ++cu[labels[i]];
    });
    for (int64_t c = 0; c < nc; ++c) {
        int64_t run = 0;
        for (int u = 0; u < nth; ++u) {
            int64_t& p = pos[(size_t)u * nc + c];
            const int64_t k = p;
            p = run;
            run += k;
        }
        members[c].resize(run);
    }
    on_threads([&](int u, int64_t i0, int64_t i1) {
        int64_t* pu = pos.data() + (size_t)u * nc;
        for (int64_t i = i0; i < i1; ++i) {
            const int64_t c = labels[i];
            members[c][pu[c]++] = i;
        }
    });
}

struct DeviceGraph {
    si::StreamBuf<int64_t> ip;
    si::StreamBuf<int32_t> ix;
    si::StreamBuf<float> w;
};

void upload_graph(DeviceGraph& g, int64_t n, const int64_t* ip, const int32_t* ix, const float* w,
                  hipStream_t s) {
    const int64_t nnz = ip[n];
    g.ip.alloc(n + 1, s);
    g.ix.alloc(nnz, s);
    g.w.alloc(nnz, s);
    si::upload_to_device(g.ip.get(), ip, sizeof(int64_t) * (n + 1), s);
    if (nnz > 0) {
        si::upload_to_device(g.ix.get(), ix, sizeof(int32_t) * nnz, s);
        si::upload_to_device(g.w.get(), w, sizeof(float) * nnz, s);
    }
    SI_HIP(hipStreamSynchronize(s));
}

// SpMM shape autotuning for graphs large enough to repay it (~10 launches per
// width), once per graph shape and context
void tune_once(si::Context& c, si::DevLaplacian& L, int widths) {
    widths &= ~L.tuned;
    if (!widths || L.nnz < (int64_t)1 << 20) return;
    const si::TuneKey key{L.n, L.nnz, L.n_long, L.max_row_len, L.order};
    si::TunedShapes& t = c.tune_cache[key];
    if (t.tuned & widths & si::kTuned16) L.shape16 = t.shape16;
    if (t.tuned & widths & si::kTuned32) L.shape32 = t.shape32;
    L.tuned |= t.tuned & widths;
    widths &= ~t.tuned;
    if (!widths) return;
    si::spmm_autotune(L, c.stream, widths);
    if (widths & si::kTuned16) t.shape16 = L.shape16;
    if (widths & si::kTuned32) t.shape32 = L.shape32;
    t.tuned |= widths;
}

void check_host_csr(int64_t n, const int64_t* ip) {
    SI_CHECK(ip != nullptr, si::kErrInvalidArgument, "indptr is null");
    SI_CHECK(ip[0] == 0, si::kErrInvalidGraph, "invalid graph: indptr[0] must be 0");
    for (int64_t i = 0; i < n; ++i)
        SI_CHECK(ip[i + 1] >= ip[i], si::kErrInvalidGraph, "invalid graph: indptr must be non-decreasing");
}

// Upload a host graph and build its operator on the device (the build and
// the SpMM autotuning count as Laplacian time).
void build_operator(si::Context& c, si::DevLaplacian& L, int64_t n, const int64_t* ip,
                    const int32_t* ix, const float* w, int mode, bool reorder, int tune_widths,
                    double* lap_ms) {
    hipStream_t s = c.stream;
    DeviceGraph g;
    const auto t0 = std::chrono::steady_clock::now();
    upload_graph(g, n, ip, ix, w, s);
    if (si::trace_on())
        std::fprintf(stderr, "[spectral_init] %-22s %8.1f ms\n", "graph upload",
                     std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    si::EventTimer t(s);
    si::laplacian_build_device(L, n, g.ip, g.ix, g.w, ip[n], mode, s, reorder, ip, ix);
    tune_once(c, L, tune_widths);
    const double ms = t.stop_ms();
    if (lap_ms) *lap_ms += ms;
}

// Run the chain on a built operator of a connected graph; host eigenpairs in
// the caller's node order.
// With `keep`, the eigenvector block stays on the device (in keep) and vec is
// left empty.
int solve_operator(si::Context& c, const si::DevLaplacian& L, int n_components, uint64_t seed,
                   const si::ChainOptions& opt, std::vector<double>& ev, std::vector<double>& vec,
                   si::SolveStats& st, double* solve_ms, si::DevBuf<double>* keep = nullptr) {
    hipStream_t s = c.stream;
    const int64_t n = L.n;
    const int k = n_components + 1;
    ev.assign(k, 0.0);
    si::DevBuf<double> dv((size_t)n * k), dx((size_t)n * k);
    si::EventTimer t(s);
    const int level = si::solve_eigenproblem_device(L, n_components, seed, opt, ev.data(), dv, st, s);
    si::rows_to_external(L, dv, k, dx, k, k, s);
    const double ms = t.stop_ms();
    if (solve_ms) *solve_ms += ms;
    if (keep) {
        *keep = std::move(dx);
        vec.clear();
    } else {
        vec.resize((size_t)n * k);
        si::to_host(vec.data(), dx.get(), vec.size(), s);
    }
    return level;
}

// The columns of an eigenvector block in selection order: ascending
// eigenvalue (stable), the trivial first one skipped (selection.rs:11-29).
std::vector<int32_t> selection_columns(const std::vector<double>& ev, int d) {
    std::vector<int> idx(ev.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return ev[a] < ev[b]; });
    std::vector<int32_t> cols(d);
    for (int c = 0; c < d; ++c) cols[c] = idx[1 + c];
    return cols;
}

// scale_and_add_noise (src/scaling.rs:78-89) of host coordinates: max-abs
// 10, plus f32 N(0, 1e-4) noise (Philox stream kStreamNoise, index = the
// flat output position, drawn on the device).
void scale_and_add_noise(si::Context& c, const std::vector<double>& coords, int d, uint64_t seed,
                         float* out) {
    double mx = 0.0;
    for (double v : coords) mx = std::max(mx, std::fabs(v));
    if (!(mx > 0.0))
        throw si::Error(si::kErrInvalidGraph,
                        "invalid graph: scale_coords: embedding is all-zero; cannot scale to max=10");
    const double ex = 10.0 / mx;
    hipStream_t s = c.stream;
    const size_t cnt = coords.size();
    double* nz = si::ws_get<double>(si::kWsAbiOut, cnt);
    si::fill_normal_f64(nz, (int64_t)(cnt / d), d, d, seed, si::kStreamNoise, s);
    std::vector<double> hn(cnt);
    si::to_host(hn.data(), nz, cnt, s);
    for (size_t i = 0; i < cnt; ++i) {
        const float noise = (float)(1e-4 * hn[i]);
        out[i] = (float)(coords[i] * ex) + noise;
    }
}

// Frees a device operator when the scope ends.
struct OperatorScope {
    si::DevLaplacian L;
    ~OperatorScope() { si::laplacian_free(L); }
};

// Solve one connected graph: build L, run the chain, return host eigenpairs.
int solve_graph(si::Context& c, int64_t n, const int64_t* ip, const int32_t* ix, const float* w,
                int n_components, uint64_t seed, int mode, const si::ChainOptions& opt,
                std::vector<double>& ev, std::vector<double>& vec, si::SolveStats& st,
                double* lap_ms, double* solve_ms) {
    OperatorScope op;
    build_operator(c, op.L, n, ip, ix, w, mode, /*reorder=*/n >= 4096,
                   opt.solver == SI_SOLVER_FAST
                       ? si::fast_tune_widths(n, n_components + 1, opt.fast.block_size)
                       : 0,
                   lap_ms);
    return solve_operator(c, op.L, n_components, seed, opt, ev, vec, st, solve_ms);
}

// Solve a Laplacian given by FP64 off-diagonal values and degrees (the
// centroid graph of the spectral meta-embedding); host eigenpairs out.
void solve_values(si::Context& c, int64_t n, const int64_t* ip, const int32_t* ix,
                  const double* lval, const double* deg, int n_components, uint64_t seed,
                  const si::ChainOptions& opt, std::vector<double>& ev, std::vector<double>& vec,
                  si::SolveStats& st) {
    hipStream_t s = c.stream;
    si::DevLaplacian L;
    si::laplacian_from_values_device(L, n, ip, ix, lval, deg, s);
    const int k = n_components + 1;
    ev.assign(k, 0.0);
    try {
        si::DevBuf<double> dv((size_t)n * k);
        si::solve_eigenproblem_device(L, n_components, seed, opt, ev.data(), dv, st, s);
        vec.resize((size_t)n * k);
        si::to_host(vec.data(), dv.get(), vec.size(), s);
    } catch (...) {
        si::laplacian_free(L);
        throw;
    }
    si::laplacian_free(L);
}

}  // namespace

namespace si {
// batched-component solves: capture FP32 rounds as hipGraphs (a test seam can
// turn it off) and the statistics of the last batched solve
bool g_batch_graphs = true;
BatchStats g_last_batch;
// spectral_init's output stage on the device (a test seam can select the host
// form, which computes the same values)
bool g_device_output = true;
}  // namespace si

namespace {

// All components of at least 2d nodes solved in one device pass: one Laplacian
// of the whole (block-diagonal) graph in locality order, where every component
// is a contiguous row range; components of <= 64 nodes by a batched dense EVD
// (one workgroup each, the reference's level 0), the others by the batched
// FP32-bulk solver.  A component the batch leaves above tol stays unsolved here
// and goes through the per-component escalation chain.
void solve_components_batched(si::Context& cx, const si::DevLaplacian& L,
                              const std::vector<int64_t>& labels,
                              const std::vector<std::vector<int64_t>>& members, int d,
                              uint64_t seed, const si::ChainOptions& opt,
                              std::vector<std::vector<double>>& cev, si::DevBuf<double>& blk,
                              std::vector<int64_t>& seg_row0, std::vector<int>& csolved,
                              si::SolveStats& last, double* solve_ms) {
    hipStream_t s = cx.stream;
    const int k = d + 1;
    const int64_t n = L.n;
    const int64_t nc = (int64_t)members.size();
    auto tp = std::chrono::steady_clock::now();
    auto phase = [&](const char* what) {  // SI_TRACE=1: phase timings
        if (!si::trace_on()) return;
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[batched] %-26s %8.1f ms\n", what,
                     std::chrono::duration<double, std::milli>(now - tp).count());
        tp = now;
    };
    {
        // row range of every component (the locality order keeps them contiguous)
        std::vector<int64_t> first(nc, -1), last_row(nc, -1);
        if (L.perm && L.comp && L.n_comp == nc) {
            // device labels of the operator's build: one pass over the rows there
            si::DevBuf<int32_t> rr(2 * nc);
            si::component_rows(n, L.perm, L.comp, nc, rr.get(), s);
            std::vector<int32_t> h(2 * nc);
            si::to_host(h.data(), rr.get(), 2 * nc, s);
            for (int64_t c = 0; c < nc; ++c)
                if (h[nc + c] >= 0) {
                    first[c] = h[c];
                    last_row[c] = h[nc + c];
                }
        } else {
            // the host labels (a build without device labels)
            SI_CHECK((int64_t)labels.size() == n, si::kErrInternal,
                     "batched components: host labels missing");
            std::vector<int32_t> perm(n);
            if (L.perm) si::to_host(perm.data(), L.perm, n, s);
            else std::iota(perm.begin(), perm.end(), 0);
            for (int64_t r = 0; r < n; ++r) {
                const int64_t c = labels[perm[r]];
                if (first[c] < 0) first[c] = r;
                last_row[c] = r;
            }
        }
        for (int64_t c = 0; c < nc; ++c)
            SI_CHECK(last_row[c] - first[c] + 1 == (int64_t)members[c].size(), si::kErrInternal,
                     "batched components: a component is not a contiguous row range");
        std::vector<int64_t> dcomp, drow0, icomp, irow0, in;
        std::vector<int32_t> dn;
        for (int64_t c = 0; c < nc; ++c) {
            const int64_t sz = (int64_t)members[c].size();
            if (sz < 2 * d) continue;
            if (sz <= si::kSmallMax) {
                dcomp.push_back(c); drow0.push_back(first[c]); dn.push_back((int32_t)sz);
            } else {
                icomp.push_back(c); irow0.push_back(first[c]); in.push_back(sz);
            }
        }
        phase("segment ranges");
        si::DevBuf<double> block((size_t)n * k);
        // rows of components too small to solve (< 2d nodes) are not written
        // by either solver
        SI_HIP(hipMemsetAsync(block.get(), 0, sizeof(double) * (size_t)n * k, s));
        std::vector<double> dev_h(dcomp.size() * k), iev(icomp.size() * k), resid;
        std::vector<int> ok;
        si::BatchStats bst;
        {
            si::EventTimer t(s);
            // The dense EVD of the small components runs on the auxiliary
            // stream, overlapping the iterative solve (each writes only the
            // rows of its own components); the main stream waits for it.
            si::DevBuf<int64_t> r0(std::max<size_t>(drow0.size(), 1));
            si::DevBuf<int32_t> nn(std::max<size_t>(dn.size(), 1));
            si::DevBuf<double> ev(std::max<size_t>(dcomp.size() * k, 1));
            hipEvent_t ready = nullptr, dense_done = nullptr;
            try {
                if (!dcomp.empty()) {
                    hipStream_t a = si::aux_stream(cx);
                    SI_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
                    SI_HIP(hipEventCreateWithFlags(&dense_done, hipEventDisableTiming));
                    SI_HIP(hipEventRecord(ready, s));  // the operator and the block exist
                    SI_HIP(hipStreamWaitEvent(a, ready, 0));
                    si::to_device(r0.get(), drow0.data(), drow0.size(), a);
                    si::to_device(nn.get(), dn.data(), dn.size(), a);
                    si::dense_seg_eig(L, (int)dcomp.size(), r0, nn, k, ev, block, k, a);
                    SI_HIP(hipEventRecord(dense_done, a));
                }
                if (!icomp.empty()) {
                    si::FastOptions fo = opt.fast;
                    fo.tol = opt.tol;
                    si::fast_solve_segments(L, irow0, in, d, seed, fo, si::g_batch_graphs,
                                            iev.data(), block, resid, ok, bst, s);
                    si::g_last_batch = bst;
                }
                if (dense_done) {
                    SI_HIP(hipStreamWaitEvent(s, dense_done, 0));
                    si::to_host(dev_h.data(), ev.get(), dev_h.size(), s);
                }
            } catch (...) {
                if (dense_done) hipStreamSynchronize(si::aux_stream(cx));
                if (ready) hipEventDestroy(ready);
                if (dense_done) hipEventDestroy(dense_done);
                throw;
            }
            if (ready) hipEventDestroy(ready);
            if (dense_done) hipEventDestroy(dense_done);
            *solve_ms += t.stop_ms();
        }
        phase("device solve");
        // the eigenvector blocks stay in the operator's row order, where every
        // component is the row range starting at seg_row0[c]
        seg_row0 = first;
        auto take = [&](int64_t c, const double* evc) {
            cev[c].assign(evc, evc + k);
            csolved[c] = 1;
        };
        for (size_t i = 0; i < dcomp.size(); ++i) {
            take(dcomp[i], dev_h.data() + i * k);
            last.level = std::max(last.level, 0);
        }
        for (size_t i = 0; i < icomp.size(); ++i) {
            if (!ok[i]) continue;
            take(icomp[i], iev.data() + i * k);
            last.level = 1;
            last.max_residual = std::max(last.max_residual, resid[i]);
        }
        last.iterations += bst.rounds;
        last.spmm_calls += bst.spmm_calls;
        blk = std::move(block);
        phase("eigenvector block");
    }
}

}  // namespace

extern "C" {

const char* si_version(void) { return "spectral-mi355x 0.1 (gfx950)"; }

void si_config_default(si_config* cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->struct_size = sizeof(si_config);
    cfg->compute_mode = SI_COMPUTE_PYTHON_COMPAT;
    cfg->solver = SI_SOLVER_FAST;
    cfg->block_size = 0;
    cfg->cheb_degree = 0;
    cfg->max_outer = 60;
    cfg->tol = 1e-5;
    cfg->dense_threshold = 2000;
    cfg->warm_start = nullptr;
    cfg->batch_components = 1;
}

si_status si_context_create(int device, si_context** out) {
    if (!out) return SI_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    si_context* ctx = nullptr;
    try {
        ctx = new si_context();
    } catch (...) {
        return SI_ERR_INTERNAL;
    }
    ctx->c.device = device;
    si_status st = guard(nullptr, [&] {
        int count = 0;
        SI_HIP(hipGetDeviceCount(&count));
        SI_CHECK(device >= 0 && device < count, si::kErrInvalidArgument, "no such HIP device");
        SI_HIP(hipSetDevice(device));
        SI_HIP(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        SI_HIP(hipGetDeviceProperties(&prop, device));
        ctx->c.num_cus = prop.multiProcessorCount;
        // the allocation-free residual's static scratch: its address on this device
        si::residual_noalloc_setup(device);
    });
    if (st != SI_OK) {
        delete ctx;
        return st;
    }
    *out = ctx;
    return SI_OK;
}

void si_context_destroy(si_context* ctx) {
    if (!ctx) return;
    if (ctx->c.stream || ctx->c.aux) hipSetDevice(ctx->c.device);
    if (ctx->c.aux) {
        hipStreamSynchronize(ctx->c.aux);
        hipStreamDestroy(ctx->c.aux);
    }
    if (ctx->c.stream) {
        hipStreamSynchronize(ctx->c.stream);
        hipStreamDestroy(ctx->c.stream);
    }
    si::spmm_profile_release(&ctx->c.profile);
    delete ctx;
}

const char* si_last_error(const si_context* ctx) {
    return ctx ? ctx->c.last_error.c_str() : "null context";
}

si_status si_laplacian_build_ex(si_context* ctx, int64_t n, const int64_t* indptr,
                                const int32_t* indices, const float* data, int input_on_device,
                                int32_t compute_mode, uint32_t flags, si_laplacian** out) {
    if (!ctx || !out) return SI_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    return guard(ctx, [&] {
        SI_CHECK(n > 0, si::kErrInvalidGraph, "invalid graph: n must be > 0");
        SI_CHECK(compute_mode == 0 || compute_mode == 1, si::kErrInvalidArgument, "unknown compute mode");
        const bool reorder = (flags & (SI_BUILD_LOCALITY_ORDER | SI_BUILD_RCM_ORDER)) != 0;
        const bool prefer_rcm = (flags & SI_BUILD_RCM_ORDER) != 0;
        hipStream_t s = ctx->c.stream;
        std::unique_ptr<si_laplacian> lap(new si_laplacian());
        lap->device = ctx->c.device;
        if (input_on_device) {
            int64_t nnz = 0;
            SI_HIP(hipMemcpyAsync(&nnz, indptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
            SI_HIP(hipStreamSynchronize(s));
            si::laplacian_build_device(lap->L, n, indptr, indices, data, nnz, compute_mode, s, reorder,
                                       nullptr, nullptr, prefer_rcm);
        } else {
            check_host_csr(n, indptr);
            DeviceGraph g;
            upload_graph(g, n, indptr, indices, data, s);
            si::laplacian_build_device(lap->L, n, g.ip, g.ix, g.w, indptr[n], compute_mode, s,
                                       reorder, indptr, indices, prefer_rcm);
        }
        *out = lap.release();
    });
}

si_status si_laplacian_build(si_context* ctx, int64_t n, const int64_t* indptr,
                             const int32_t* indices, const float* data, int input_on_device,
                             int32_t compute_mode, si_laplacian** out) {
    const uint32_t flags = n >= 4096 ? SI_BUILD_LOCALITY_ORDER : 0u;
    return si_laplacian_build_ex(ctx, n, indptr, indices, data, input_on_device, compute_mode,
                                 flags, out);
}

void si_laplacian_destroy(si_laplacian* lap) {
    if (!lap) return;
    hipSetDevice(lap->device);
    si::laplacian_free(lap->L);
    delete lap;
}

si_status si_laplacian_info(const si_laplacian* lap, int64_t* n, int64_t* nnz, int64_t* n_long,
                            int32_t* degree_exact) {
    if (!lap) return SI_ERR_INVALID_ARGUMENT;
    if (n) *n = lap->L.n;
    if (nnz) *nnz = lap->L.nnz;
    if (n_long) *n_long = lap->L.n_long;
    if (degree_exact) *degree_exact = lap->L.degree_exact;
    return SI_OK;
}

si_status si_laplacian_export(si_context* ctx, const si_laplacian* lap, double* degrees,
                              double* sqrt_deg, int64_t* l_indptr, int32_t* l_indices,
                              double* l_data, int64_t* nnz_l) {
    if (!ctx || !lap) return SI_ERR_INVALID_ARGUMENT;
    return guard(ctx, [&] {
        hipStream_t s = ctx->c.stream;
        const si::DevLaplacian& L = lap->L;
        const int64_t n = L.n;
        std::vector<int32_t> perm(n), iperm(n);
        if (L.perm) {
            si::to_host(perm.data(), L.perm, n, s);
            si::to_host(iperm.data(), L.iperm, n, s);
        } else {
            for (int64_t i = 0; i < n; ++i) perm[i] = iperm[i] = (int32_t)i;
        }
        std::vector<double> tmp(n);
        if (degrees) {
            si::to_host(tmp.data(), L.deg, n, s);
            for (int64_t v = 0; v < n; ++v) degrees[v] = tmp[iperm[v]];
        }
        if (sqrt_deg) {
            si::to_host(tmp.data(), L.sqrt_deg, n, s);
            for (int64_t v = 0; v < n; ++v) sqrt_deg[v] = tmp[iperm[v]];
        }
        std::vector<int32_t> rp(n + 1), colp(L.nnz);
        std::vector<double> valp(L.nnz);
        si::to_host(rp.data(), L.rowptr, n + 1, s);
        if (L.nnz) {
            si::to_host(colp.data(), L.col, L.nnz, s);
            si::to_host(valp.data(), L.l64, L.nnz, s);
        }
        // reference layout in the caller's node order: off-diagonal entries plus
        // the unit diagonal, ascending columns
        std::vector<std::pair<int32_t, double>> row;
        int64_t o = 0;
        if (l_indptr) l_indptr[0] = 0;
        for (int64_t i = 0; i < n; ++i) {
            const int32_t ii = iperm[i];
            row.clear();
            for (int32_t p = rp[ii]; p < rp[ii + 1]; ++p) row.emplace_back(perm[colp[p]], valp[p]);
            std::sort(row.begin(), row.end(),
                      [](const std::pair<int32_t, double>& x, const std::pair<int32_t, double>& y) {
                          return x.first < y.first;
                      });
            bool diag_done = false;
            for (const auto& e : row) {
                const int32_t j = e.first;
                if (j == i) continue;
                if (!diag_done && j > i) {
                    if (l_indices) l_indices[o] = (int32_t)i;
                    if (l_data) l_data[o] = 1.0;
                    ++o;
                    diag_done = true;
                }
                if (l_indices) l_indices[o] = j;
                if (l_data) l_data[o] = e.second;
                ++o;
            }
            if (!diag_done) {
                if (l_indices) l_indices[o] = (int32_t)i;
                if (l_data) l_data[o] = 1.0;
                ++o;
            }
            if (l_indptr) l_indptr[i + 1] = o;
        }
        if (nnz_l) *nnz_l = o;
    });
}

si_status si_spmm(si_context* ctx, const si_laplacian* lap, int32_t precision, const void* x,
                  void* y, int32_t k, int32_t ld) {
    if (!ctx || !lap || !x || !y) return SI_ERR_INVALID_ARGUMENT;
    return guard(ctx, [&] {
        SI_CHECK(si::spmm_width_supported(k), si::kErrInvalidArgument, "spmm: unsupported width");
        SI_CHECK(ld >= k, si::kErrInvalidArgument, "spmm: ld < k");
        const si::DevLaplacian& L = lap->L;
        hipStream_t s = ctx->c.stream;
        const size_t cnt = (size_t)L.n * k;
        if (precision == SI_F32) {
            if (!L.perm) {
                si::spmm_f32(L, (const float*)x, (float*)y, k, ld, 1.f, 0.f, 0.f, nullptr, s);
            } else {
                si::DevBuf<float> xi(cnt), yi(cnt);
                si::rows_to_internal(L, (const float*)x, ld, xi.get(), k, k, s);
                si::spmm_f32(L, xi, yi, k, k, 1.f, 0.f, 0.f, nullptr, s);
                si::rows_to_external(L, yi.get(), k, (float*)y, ld, k, s);
                SI_HIP(hipStreamSynchronize(s));
            }
        } else {
            if (!L.perm) {
                si::spmm_f64(L, (const double*)x, (double*)y, k, ld, 1.0, 0.0, 0.0, nullptr, s);
            } else {
                si::DevBuf<double> xi(cnt), yi(cnt);
                si::rows_to_internal(L, (const double*)x, ld, xi.get(), k, k, s);
                si::spmm_f64(L, xi, yi, k, k, 1.0, 0.0, 0.0, nullptr, s);
                si::rows_to_external(L, yi.get(), k, (double*)y, ld, k, s);
                SI_HIP(hipStreamSynchronize(s));
            }
        }
        SI_HIP(hipStreamSynchronize(s));
    });
}

si_status si_chebyshev_filter(si_context* ctx, const si_laplacian* lap, const double* r,
                              int32_t k, double a, double b, int32_t degree, double* out) {
    if (!ctx || !lap || !r || !out) return SI_ERR_INVALID_ARGUMENT;
    return guard(ctx, [&] {
        SI_CHECK(si::spmm_width_supported(k), si::kErrInvalidArgument, "chebyshev: unsupported width");
        SI_CHECK(b > a, si::kErrInvalidArgument, "chebyshev: need b > a");
        const si::DevLaplacian& L = lap->L;
        hipStream_t s = ctx->c.stream;
        si::DevBuf<double> ri((size_t)L.n * k), oi((size_t)L.n * k);
        si::rows_to_internal(L, r, k, ri.get(), k, k, s);
        si::chebyshev_filter_f64(L, ri, k, a, b, degree, oi, s);
        si::rows_to_external(L, oi.get(), k, out, k, k, s);
        SI_HIP(hipStreamSynchronize(s));
    });
}

si_status si_operator_apply_host(si_context* ctx, const si_laplacian* lap, const double* x,
                                 double* y, int32_t k) {
    if (!ctx || !lap || !x || !y) return SI_ERR_INVALID_ARGUMENT;
    return guard(ctx, [&] {
        SI_CHECK(si::spmm_width_supported(k), si::kErrInvalidArgument,
                 "operator apply: need 1 <= k <= 4096");
        const si::DevLaplacian& L = lap->L;
        hipStream_t s = ctx->c.stream;
        const size_t nk = (size_t)L.n * k;
        si::DevBuf<double> xe(nk), xi(nk), yi(nk);
        si::to_device(xe.get(), x, nk, s);
        si::rows_to_internal(L, xe.get(), k, xi.get(), k, k, s);
        si::spmm_f64(L, xi, yi, k, k, 1.0, 0.0, 0.0, nullptr, s);
        si::rows_to_external(L, yi.get(), k, xe.get(), k, k, s);
        si::to_host(y, xe.get(), nk, s);
    });
}

si_status si_spmv_csr(si_context* ctx, int64_t n_rows, const uint64_t* indptr,
                      const uint64_t* indices, const double* data, int64_t data_len,
                      const double* x, int64_t x_len, double* y) {
    if (!ctx || n_rows < 0 || x_len < 0 || data_len < 0 || !indptr) return SI_ERR_INVALID_ARGUMENT;
    return guard(ctx, [&] {
        const uint64_t nnz = indptr[n_rows];
        // the reference's preconditions (operator.rs:33-48), as errors
        SI_CHECK(data_len >= 0 && (uint64_t)data_len >= nnz, si::kErrInvalidArgument,
                 "CSR invariant violated: data.len() < nnz");
        SI_CHECK(nnz == 0 || (indices && data), si::kErrInvalidArgument, "CSR: missing arrays");
        SI_CHECK(n_rows == 0 || y, si::kErrInvalidArgument, "CSR: missing y");
        SI_CHECK(nnz < (1ull << 31) && x_len < (1ll << 31), si::kErrUnsupported,
                 "spmv_csr: more than 2^31 entries or columns");
        std::vector<int64_t> ip(n_rows + 1);
        for (int64_t i = 0; i <= n_rows; ++i) {
            SI_CHECK(indptr[i] <= nnz && (i == 0 || indptr[i] >= indptr[i - 1]),
                     si::kErrInvalidArgument, "CSR invariant violated: indptr not monotone");
            ip[i] = (int64_t)indptr[i];
        }
        std::vector<int32_t> ix(nnz);
        for (uint64_t p = 0; p < nnz; ++p) {
            SI_CHECK(indices[p] < (uint64_t)x_len, si::kErrInvalidArgument,
                     "CSR invariant violated: column index out of bounds (x.len()=" +
                         std::to_string(x_len) + ")");
            ix[p] = (int32_t)indices[p];
        }
        if (n_rows == 0) return;
        hipStream_t s = ctx->c.stream;
        si::DevBuf<int64_t> dip(n_rows + 1);
        si::DevBuf<int32_t> dix(nnz);
        si::DevBuf<double> dd(nnz), dx(x_len), dy(n_rows);
        si::to_device(dip.get(), ip.data(), ip.size(), s);
        if (nnz) {
            si::to_device(dix.get(), ix.data(), nnz, s);
            si::to_device(dd.get(), data, nnz, s);
        }
        if (x_len) si::to_device(dx.get(), x, (size_t)x_len, s);
        si::csr_spmv_exact_f64(n_rows, dip, dix, dd, dx, dy, s);
        si::to_host(y, dy.get(), (size_t)n_rows, s);
    });
}

si_status si_max_eigenpair_residual(si_context* ctx, const si_laplacian* lap,
                                    const double* evals, const double* evecs, int32_t k,
                                    double* out) {
    if (!ctx || !lap || !evals || !evecs || !out) return SI_ERR_INVALID_ARGUMENT;
    return guard(ctx, [&] {
        SI_CHECK(si::spmm_width_supported(k), si::kErrInvalidArgument, "residual: unsupported width");
        std::vector<double> res(k);
        si::DevBuf<double> vi((size_t)lap->L.n * k);
        si::rows_to_internal(lap->L, evecs, k, vi.get(), k, k, ctx->c.stream);
        si::fp64_residuals(lap->L, vi, k, evals, res.data(), ctx->c.stream);
        *out = si::max_residual_of(res.data(), k);
    });
}

si_status si_lobpcg_solve(si_context* ctx, const si_laplacian* lap, int32_t n_components,
                          uint64_t seed, int32_t regularize, double tol, double* evals,
                          double* evecs, si_report* rep) {
    if (!ctx || !lap || !evals || !evecs) return SI_ERR_INVALID_ARGUMENT;
    return guard(ctx, [&] {
        SI_CHECK(n_components >= 0 && n_components + 1 <= 64, si::kErrUnsupported,
                 "n_components above 63 is not supported");
        si::SolveStats st;
        const int k = n_components + 1;
        double* vi = si::ws_get<double>(si::kWsChainVec, (size_t)lap->L.n * k);
        si::EventTimer t(ctx->c.stream);
        const bool ok = si::lobpcg_solve_plain(lap->L, n_components, seed, regularize != 0,
                                               tol > 0 ? tol : 1e-5, evals, vi, st,
                                               ctx->c.stream);
        if (ok) {
            std::vector<double> res(k);
            si::fp64_residuals(lap->L, vi, k, evals, res.data(), ctx->c.stream);
            st.max_residual = si::max_residual_of(res.data(), k);
            si::rows_to_external(lap->L, vi, k, evecs, k, k, ctx->c.stream);
        }
        const double ms = t.stop_ms();
        st.level = 1;
        fill_report(rep, st, evals, n_components + 1, ms);
        SI_CHECK(ok, si::kErrConvergence, "lobpcg_solve produced no result");
    });
}

si_status si_lobpcg_sinv_solve(si_context* ctx, const si_laplacian* lap, int32_t n_components,
                               uint64_t seed, double* evals, double* evecs, si_report* rep) {
    if (!ctx || !lap || !evals || !evecs) return SI_ERR_INVALID_ARGUMENT;
    return guard(ctx, [&] {
        SI_CHECK(n_components >= 0 && n_components + 1 <= 64, si::kErrUnsupported,
                 "n_components above 63 is not supported");
        si::SolveStats st;
        const int k = n_components + 1;
        double* vi = si::ws_get<double>(si::kWsChainVec, (size_t)lap->L.n * k);
        si::EventTimer t(ctx->c.stream);
        const bool ok = si::lobpcg_sinv_solve_device(lap->L, n_components, seed, evals, vi, st,
                                                     ctx->c.stream);
        if (ok) {
            std::vector<double> res(k);
            si::fp64_residuals(lap->L, vi, k, evals, res.data(), ctx->c.stream);
            st.max_residual = si::max_residual_of(res.data(), k);
            si::rows_to_external(lap->L, vi, k, evecs, k, k, ctx->c.stream);
        }
        const double ms = t.stop_ms();
        st.level = 2;
        fill_report(rep, st, evals, n_components + 1, ms);
        SI_CHECK(ok, si::kErrConvergence, "lobpcg_sinv_solve produced no result");
    });
}

si_status si_solve_eigenproblem(si_context* ctx, const si_laplacian* lap, int32_t n_components,
                                uint64_t seed, const si_config* cfg, double* evals,
                                double* evecs, si_report* rep) {
    if (!ctx || !lap || !evals || !evecs) return SI_ERR_INVALID_ARGUMENT;
    return guard(ctx, [&] {
        SI_CHECK(n_components > 0 && (int64_t)n_components < lap->L.n, si::kErrTooFewNodes,
                 "graph has too few nodes for the requested embedding");
        SI_CHECK(n_components + 1 <= 64, si::kErrUnsupported, "n_components above 63 is not supported");
        si::ChainOptions o = chain_options(cfg);
        si::SolveStats st;
        const int k = n_components + 1;
        hipStream_t s = ctx->c.stream;
        // the FP32 sweeps pick their SpMM shape for this graph once (a cache
        // held by the Laplacian)
        if (o.solver == SI_SOLVER_FAST)
            tune_once(ctx->c, const_cast<si::DevLaplacian&>(lap->L),
                      si::fast_tune_widths(lap->L.n, n_components + 1, o.fast.block_size));
        const si::DevLaplacian& L = lap->L;
        double* vi = si::ws_get<double>(si::kWsChainVec, (size_t)L.n * k);
        if (o.fast.warm_start) {
            double* wsi = si::ws_get<double>(si::kWsAbiIn, (size_t)L.n * k);
            si::rows_to_internal(L, o.fast.warm_start, k, wsi, k, k, s);
            o.fast.warm_start = wsi;
        }
        si::EventTimer t(s);
        si::solve_eigenproblem_device(L, n_components, seed, o, evals, vi, st, s);
        si::rows_to_external(L, vi, k, evecs, k, k, s);
        const double ms = t.stop_ms();
        fill_report(rep, st, evals, k, ms);
    });
}

si_status si_rsvd_solve(si_context* ctx, const si_laplacian* lap, int32_t n_components,
                        uint64_t seed, int32_t oversampling, int32_t nbiter, double* evals,
                        double* evecs, si_report* rep) {
    if (!ctx || !lap || !evals || !evecs) return SI_ERR_INVALID_ARGUMENT;
    return guard(ctx, [&] {
        si::SolveStats st;
        const int k = n_components + 1;
        si::DevBuf<double> vi((size_t)lap->L.n * k);
        si::EventTimer t(ctx->c.stream);
        si::rsvd_solve_device(lap->L, n_components, seed, oversampling, nbiter, evals, vi, st,
                              ctx->c.stream);
        const double ms = t.stop_ms();
        std::vector<double> res(k);
        si::fp64_residuals(lap->L, vi, k, evals, res.data(), ctx->c.stream);
        st.max_residual = si::max_residual_of(res.data(), k);
        si::rows_to_external(lap->L, vi.get(), k, evecs, k, k, ctx->c.stream);
        SI_HIP(hipStreamSynchronize(ctx->c.stream));
        st.level = 4;
        fill_report(rep, st, evals, n_components + 1, ms);
    });
}

si_status si_dense_evd(si_context* ctx, const si_laplacian* lap, int32_t k, double* evals,
                       double* evecs) {
    if (!ctx || !lap || !evals || !evecs) return SI_ERR_INVALID_ARGUMENT;
    return guard(ctx, [&] {
        si::DevBuf<double> vi((size_t)lap->L.n * (k > 0 ? k : 1));
        si::dense_evd_device(lap->L, k, evals, vi, ctx->c.stream);
        si::rows_to_external(lap->L, vi.get(), k, evecs, k, k, ctx->c.stream);
        SI_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

si_status si_spectral_init(si_context* ctx, int64_t n, const int64_t* indptr,
                           const int32_t* indices, const float* data, int32_t n_components,
                           uint64_t seed, const float* raw_data, int64_t n_features,
                           const si_config* cfg, float* out, si_report* rep) {
    if (!ctx || !out) return SI_ERR_INVALID_ARGUMENT;
    return guard(ctx, [&] {
        // input validation (lib.rs:206-214)
        if (n == 0 || n_components == 0)
            throw si::Error(si::kErrInvalidGraph,
                            "invalid graph: spectral_init requires n > 0 and n_components > 0; got n=" +
                                std::to_string(n) + ", n_components=" + std::to_string(n_components));
        SI_CHECK(n > 0 && n_components > 0, si::kErrInvalidArgument, "negative sizes");
        if (n <= n_components)
            throw si::Error(si::kErrTooFewNodes, "graph has too few nodes (" + std::to_string(n) +
                                                     ") for " + std::to_string(n_components) +
                                                     "-dimensional embedding");
        SI_CHECK(n_components + 1 <= 64, si::kErrUnsupported, "n_components above 63 is not supported");
        check_host_csr(n, indptr);
        const int mode = cfg ? cfg->compute_mode : 0;
        si::ChainOptions opt = chain_options(cfg);
        opt.fast.warm_start = nullptr;  // warm starts are per-graph device blocks
        // batch_components is read only from callers whose struct has the field
        const bool batch = !cfg || cfg->struct_size < offsetof(si_config, batch_components) +
                                                         sizeof(cfg->batch_components) ||
                           cfg->batch_components != 0;
        const int d = n_components;
        const int k = d + 1;
        auto tp = std::chrono::steady_clock::now();
        auto phase = [&](const char* what) {  // SI_TRACE=1: pipeline phase timings
            if (!si::trace_on()) return;
            const auto now = std::chrono::steady_clock::now();
            std::fprintf(stderr, "[spectral_init] %-22s %8.1f ms\n", what,
                         std::chrono::duration<double, std::milli>(now - tp).count());
            tp = now;
        };
        // The operator of the whole graph serves the single-component solve and
        // the batched solve of every component.  For n >= 4096 it is built
        // first, in locality order, whose device form also labels the
        // components (symmetric graphs); otherwise the host sweep labels them.
        OperatorScope whole;
        double lap_ms = 0.0, solve_ms = 0.0;
        const bool large = n >= 4096;
        const bool fast = opt.solver == SI_SOLVER_FAST;
        if (large) {
            build_operator(ctx->c, whole.L, n, indptr, indices, data, mode, /*reorder=*/true,
                           /*tune_widths=*/0, &lap_ms);
            phase("operator");
        }
        std::vector<int64_t> labels;  // the host sweep's labels
        std::vector<int32_t> lab32;   // the build's device labels, read back for several components
        int64_t nc = 0;
        if (whole.L.n_comp > 0) {
            nc = whole.L.n_comp;
            if (nc > 1) {
                lab32.resize(n);
                si::to_host(lab32.data(), whole.L.comp, n, ctx->c.stream);
            }
        } else {
            nc = find_components(n, indptr, indices, labels);
        }
        phase("components");
        std::vector<double> coords;  // host coordinates (paths without the device output stage)
        bool output_done = false;
        si::SolveStats last;
        std::vector<double> ev, vec, sel;
        if (nc == 1) {
            if (!large)
                build_operator(ctx->c, whole.L, n, indptr, indices, data, mode, /*reorder=*/false,
                               /*tune_widths=*/0, &lap_ms);
            if (fast) {
                si::EventTimer t(ctx->c.stream);
                tune_once(ctx->c, whole.L, si::fast_tune_widths(n, k, opt.fast.block_size));
                lap_ms += t.stop_ms();
            }
            si::DevBuf<double> blk;
            last.level = solve_operator(ctx->c, whole.L, d, seed, opt, ev, vec, last, &solve_ms,
                                        si::g_device_output ? &blk : nullptr);
            phase("solve");
            if (!si::g_device_output) {
                select_and_sign(ev, vec, n, k, d, coords);
            } else {
                si::EmbedPlan plan;
                plan.n = n;
                plan.d = d;
                plan.b = blk.get();
                plan.ldb = k;
                plan.kind = {1};
                plan.row0 = {0};
                plan.rows = {n};
                plan.cols = selection_columns(ev, d);
                plan.seed = seed;
                plan.out = out;
                si::embed_output_device(plan, ctx->c.stream);
                output_done = true;
            }
        } else {
            // embed_disconnected (src/multi_component.rs:16-74)
            std::vector<std::vector<int64_t>> members;
            if (lab32.empty()) member_lists(labels, nc, members);
            else member_lists(lab32, nc, members);
            std::vector<double> meta((size_t)nc * d, 0.0);
            const bool spectral_meta = nc > 2 * d;
            if (spectral_meta && !raw_data)
                throw si::Error(si::kErrInvalidGraph,
                                "invalid graph: data is required for spectral meta-embedding when "
                                "n_conn_components > 2 * n_embedding_dims");
            // spectral_meta_embedding (multi_component.rs:143-220) starts with the
            // components' centroids of the raw data, summed in member order: on
            // host threads that run while the device solves the components
            CentroidJob centroids;
            if (spectral_meta) centroids.start(raw_data, n_features, members);
            // Eigenpairs of every component that gets its own embedding (size >= 2d):
            // all at once on the device (batched), or one solve per component.
            std::vector<std::vector<double>> cev(nc), cvec(nc);
            std::vector<int> csolved(nc, 0);
            si::DevBuf<double> blk;        // batch-solved eigenvectors, operator row order
            std::vector<int64_t> seg_row0;  // each component's first row in blk
            std::vector<int64_t> lookup(n, -1);
            auto solve_one = [&](int64_t c) {
                const std::vector<int64_t>& mem = members[c];
                const int64_t sz = (int64_t)mem.size();
                for (int64_t q = 0; q < sz; ++q) lookup[mem[q]] = q;
                std::vector<int64_t> sip(sz + 1, 0);
                std::vector<int32_t> six;
                std::vector<float> sw;
                for (int64_t q = 0; q < sz; ++q) {
                    const int64_t u = mem[q];
                    for (int64_t p = indptr[u]; p < indptr[u + 1]; ++p)
                        if (lookup[indices[p]] >= 0) {
                            six.push_back((int32_t)lookup[indices[p]]);
                            sw.push_back(data[p]);
                        }
                    sip[q + 1] = (int64_t)six.size();
                }
                si::SolveStats cst;
                last.level = solve_graph(ctx->c, sz, sip.data(), six.data(), sw.data(), d, seed, mode,
                                         opt, cev[c], cvec[c], cst, &lap_ms, &solve_ms);
                last.iterations += cst.iterations;
                last.spmm_calls += cst.spmm_calls;
                last.max_residual = std::max(last.max_residual, cst.max_residual);
                for (int64_t node : mem) lookup[node] = -1;
                csolved[c] = 1;
            };
            // the batched segment solver carries k = d + 1 <= 8 columns; wider
            // embeddings take the per-component loop
            bool all_batched = false;
            std::exception_ptr solve_error;
            try {
                if (!(batch && d <= 7)) si::laplacian_free(whole.L);
                if (batch && d <= 7) {
                    // the batch needs every component as a contiguous row range:
                    // the locality order
                    if (!whole.L.perm)
                        build_operator(ctx->c, whole.L, n, indptr, indices, data, mode,
                                       /*reorder=*/true, /*tune_widths=*/0, &lap_ms);
                    {
                        // the batched solver's block is 16 columns wide
                        si::EventTimer t(ctx->c.stream);
                        tune_once(ctx->c, whole.L, si::kTuned16);
                        lap_ms += t.stop_ms();
                    }
                    phase("spmm autotune");
                    solve_components_batched(ctx->c, whole.L, labels, members, d, seed, opt, cev, blk,
                                             seg_row0, csolved, last, &solve_ms);
                    phase("batched solve");
                }
                // components left to the per-component chain (a batch failure, or
                // no batch) make the placement a host step
                all_batched = blk.get() != nullptr && si::g_device_output;
                for (int64_t c = 0; c < nc; ++c)
                    if (!csolved[c] && (int64_t)members[c].size() >= 2 * d) {
                        all_batched = false;
                        solve_one(c);
                    }
            } catch (...) {
                solve_error = std::current_exception();
            }
            centroids.join();
            phase("component eigenpairs");
            // the rest of the meta-embedding; its own errors are reported
            // before a component solve's, in the reference's order of steps
            if (spectral_meta) {
                const std::vector<double>& cen = centroids.cen;
                // Gaussian affinity exp(-||c_i - c_j||^2) in FP64
                std::vector<double> aff((size_t)nc * nc);
                for (int64_t i = 0; i < nc; ++i)
                    for (int64_t j = 0; j < nc; ++j) {
                        double d2 = 0.0;
                        for (int64_t f = 0; f < n_features; ++f) {
                            const double df = cen[i * n_features + f] - cen[j * n_features + f];
                            d2 += df * df;
                        }
                        aff[i * nc + j] = std::exp(-d2);
                    }
                // normalized Laplacian of the centroid affinity graph, built in FP64
                // as the reference does (degrees over off-diagonal entries,
                // L_ij = -isd_i * aff_ij * isd_j): the affinities are often far below
                // the FP32 range, the normalised entries are O(1)
                std::vector<double> mdeg(nc, 0.0), misd(nc, 0.0);
                for (int64_t i = 0; i < nc; ++i) {
                    double sum = 0.0;
                    for (int64_t j = 0; j < nc; ++j)
                        if (j != i) sum += aff[i * nc + j];
                    mdeg[i] = sum;
                    misd[i] = sum > 0.0 ? 1.0 / std::sqrt(sum) : 0.0;
                }
                std::vector<int64_t> mip(nc + 1);
                std::vector<int32_t> mix;
                std::vector<double> mval;
                mip[0] = 0;
                for (int64_t i = 0; i < nc; ++i) {
                    for (int64_t j = 0; j < nc; ++j)
                        if (j != i) {
                            mix.push_back((int32_t)j);
                            mval.push_back(-misd[i] * aff[i * nc + j] * misd[j]);
                        }
                    mip[i + 1] = (int64_t)mix.size();
                }
                si::ChainOptions mo = opt;
                si::SolveStats mst;
                solve_values(ctx->c, nc, mip.data(), mix.data(), mval.data(), mdeg.data(), d, seed,
                             mo, ev, vec, mst);
                std::vector<int> idx(k);
                std::iota(idx.begin(), idx.end(), 0);
                std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return ev[a] < ev[b]; });
                double mx = 0.0;
                for (int64_t r = 0; r < nc; ++r)
                    for (int c = 0; c < d; ++c) {
                        meta[r * d + c] = vec[(size_t)r * k + idx[1 + c]];
                        mx = std::max(mx, std::fabs(meta[r * d + c]));
                    }
                if (mx == 0.0)
                    throw si::Error(si::kErrInvalidGraph, "invalid graph: spectral meta-embedding produced all-zero result");
                for (double& v : meta) v /= mx;
                // eigenvector signs are the eigensolver's choice (faer's in the
                // reference): fix them as normalize_signs does, so that the
                // placement does not depend on the solver's internals
                for (int c = 0; c < d; ++c) {
                    double best = meta[c];
                    for (int64_t r = 1; r < nc; ++r)
                        if (std::fabs(meta[r * d + c]) > std::fabs(best)) best = meta[r * d + c];
                    if (std::signbit(best))
                        for (int64_t r = 0; r < nc; ++r) meta[r * d + c] = -meta[r * d + c];
                }
            } else {
                // orthogonal_meta_embedding (multi_component.rs:127-139)
                for (int64_t r = 0; r < nc; ++r) meta[r * d + r % d] = r < d ? 1.0 : -1.0;
            }
            // compute_data_range (multi_component.rs:224-255)
            std::vector<double> range(nc, 1.0);
            for (int64_t c = 0; c < nc && nc > 1; ++c) {
                double md = INFINITY;
                for (int64_t i = 0; i < nc; ++i) {
                    if (i == c) continue;
                    double s2 = 0.0;
                    for (int q = 0; q < d; ++q) {
                        const double df = meta[c * d + q] - meta[i * d + q];
                        s2 += df * df;
                    }
                    const double dist = std::sqrt(s2);
                    if (dist > 0.0 && dist < md) md = dist;
                }
                range[c] = std::isfinite(md) ? md / 2.0 : 1.0;
            }
            phase("meta-embedding, ranges");
            if (solve_error) std::rethrow_exception(solve_error);
            for (int64_t c = 0; c < nc; ++c)
                if (!cev[c].empty()) ev = cev[c];
            if (all_batched) {
                // every component's placement, the scaling and the noise on the
                // device, from the batch's block
                si::DevBuf<int32_t> dlab;
                const int32_t* lab = whole.L.comp;
                if (!whole.L.n_comp) {
                    std::vector<int32_t> l32(labels.begin(), labels.end());
                    dlab.alloc(n);
                    si::to_device(dlab.get(), l32.data(), n, ctx->c.stream);
                    lab = dlab.get();
                }
                si::EmbedPlan plan;
                plan.n = n;
                plan.d = d;
                plan.b = blk.get();
                plan.ldb = k;
                plan.rowmap = whole.L.iperm;
                plan.rownode = whole.L.perm;
                plan.lab = lab;
                plan.nseg = (int)nc;
                plan.kind.assign(nc, 0);
                plan.cols.assign((size_t)nc * d, 0);
                plan.row0.assign(nc, 0);
                plan.rows.assign(nc, 0);
                for (int64_t c = 0; c < nc; ++c) {
                    if ((int64_t)members[c].size() < 2 * d) continue;
                    plan.kind[c] = 1;
                    plan.row0[c] = seg_row0[c];
                    plan.rows[c] = (int64_t)members[c].size();
                    const std::vector<int32_t> cc = selection_columns(cev[c], d);
                    std::copy(cc.begin(), cc.end(), plan.cols.begin() + c * d);
                }
                plan.affine = true;
                plan.range = range;
                plan.meta = meta;
                plan.seed = seed;
                plan.out = out;
                si::embed_output_device(plan, ctx->c.stream);
                output_done = true;
            } else if (blk.get()) {
                // batch-solved components in member order, for the host placement
                std::vector<double> hb((size_t)n * k);
                std::vector<int32_t> iperm(n);
                si::to_host(hb.data(), blk.get(), hb.size(), ctx->c.stream);
                si::to_host(iperm.data(), whole.L.iperm, n, ctx->c.stream);
                for (int64_t c = 0; c < nc; ++c) {
                    if (!csolved[c] || !cvec[c].empty()) continue;
                    const std::vector<int64_t>& mem = members[c];
                    cvec[c].resize(mem.size() * k);
                    for (size_t q = 0; q < mem.size(); ++q) {
                        const size_t row = (size_t)iperm[mem[q]];
                        for (int j = 0; j < k; ++j) cvec[c][q * k + j] = hb[row * k + j];
                    }
                }
            }
            // embed_single_component (multi_component.rs:261-312): place each component
            if (!output_done) coords.assign((size_t)n * d, 0.0);
            for (int64_t c = 0; c < nc && !output_done; ++c) {
                const std::vector<int64_t>& mem = members[c];
                const int64_t sz = (int64_t)mem.size();
                if (sz < 2 * d) {
                    for (int64_t node : mem)
                        for (int q = 0; q < d; ++q) coords[node * d + q] = meta[c * d + q];
                    continue;
                }
                select_and_sign(cev[c], cvec[c], sz, k, d, sel);
                double mx = 0.0;
                for (double v : sel) mx = std::max(mx, std::fabs(v));
                const double f = mx > 0.0 ? range[c] / mx : 1.0;
                for (int64_t q = 0; q < sz; ++q)
                    for (int e = 0; e < d; ++e)
                        coords[mem[q] * d + e] = sel[(size_t)q * d + e] * f + meta[c * d + e];
            }
        }
        phase("select and place");
        if (!output_done) scale_and_add_noise(ctx->c, coords, d, seed, out);
        phase("scale and noise");
        if (rep) {
            fill_report(rep, last, ev.empty() ? nullptr : ev.data(), ev.empty() ? 0 : (int)ev.size(), solve_ms);
            rep->laplacian_ms = lap_ms;
        }
    });
}

}  // extern "C"
