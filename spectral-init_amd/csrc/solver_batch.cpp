// solver_batch.cpp -- the connected components of one graph solved together.
//
// The reference embeds a disconnected graph one component at a time
// (embed_disconnected -> embed_single_component -> solve_eigenproblem,
// src/multi_component.rs:16-123, 261-312).  Here the components stay in one
// block-diagonal device Laplacian whose locality order keeps every component a
// contiguous row segment, and one FP32-bulk Chebyshev block solve serves all of
// them at once:
//
//   * every filter step is one SpMM launch over all rows; each row takes the
//     Chebyshev scalars of its segment from a small table, so every segment has
//     its own damped interval [a_c, 2] and degree (a segment past its degree, or
//     converged, copies through without gathering);
//   * the Rayleigh-Ritz Gram matrices, the m x m generalized eigenproblems and
//     the X = Y C / AX = LY C update are segmented kernels (one workgroup per
//     segment for the small problems);
//   * each FP32 round (filter + Rayleigh-Ritz) is captured once per filter
//     degree as a hipGraph and replayed, so a round is one graph launch plus one
//     small upload and one small download.
//
// The FP64 finish (Rayleigh-Ritz on the lowest mf Ritz vectors and the FP64
// residual against the exact Laplacian) is segmented the same way.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>

#include "si_host.hpp"

namespace si {

namespace {

struct WorkList {
    DevBuf<int64_t> rows;
    DevBuf<int32_t> seg;
    DevBuf<int32_t> segblk;
    SegWork w;
};

// Split every segment into items of at most `chunk` rows.
void build_worklist(const std::vector<int64_t>& row0, const std::vector<int64_t>& rown,
                    int64_t chunk, WorkList& wl, hipStream_t s) {
    std::vector<int64_t> hr;
    std::vector<int32_t> hs, hb(row0.size() + 1, 0);
    for (size_t c = 0; c < row0.size(); ++c) {
        hb[c] = (int32_t)hs.size();
        for (int64_t r = 0; r < rown[c]; r += chunk) {
            hr.push_back(row0[c] + r);
            hr.push_back(row0[c] + std::min(rown[c], r + chunk));
            hs.push_back((int32_t)c);
        }
    }
    hb[row0.size()] = (int32_t)hs.size();
    wl.rows.alloc(hr.size());
    wl.seg.alloc(hs.size());
    wl.segblk.alloc(hb.size());
    to_device(wl.rows.get(), hr.data(), hr.size(), s);
    to_device(wl.seg.get(), hs.data(), hs.size(), s);
    to_device(wl.segblk.get(), hb.data(), hb.size(), s);
    wl.w.nseg = (int)row0.size();
    wl.w.nblk = (int64_t)hs.size();
    wl.w.blk_rows = wl.rows.get();
    wl.w.blk_seg = wl.seg.get();
    wl.w.seg_blk = wl.segblk.get();
}

// Largest degree d <= cap with T_d(c/e) <= e^9.2 on [lo, b] (the
// amplification cap of the single-graph solver).
int capped_degree(double lo, double b, int cap) {
    const double c = (lo + b) / 2.0, e = (b - lo) / 2.0;
    const double x0 = c / e;
    const int dmax = (int)std::floor(9.2 / std::acosh(std::max(x0, 1.0 + 1e-12)));
    return std::max(2, std::min(cap, dmax));
}

// Per-step, per-segment epilogue scalars of a filter of per-segment degree
// deg[c] (0 = frozen) on [a[c], b]; row `nseg` of each step is the frozen
// segment of rows that belong to no segment.
template <typename T>
void coef_table(const std::vector<double>& a, const std::vector<int>& deg, double b, int steps,
                std::vector<T>& tab) {
    const int nseg = (int)a.size();
    tab.assign((size_t)steps * (nseg + 1) * 3, T(0));
    for (int j = 1; j <= steps; ++j) {
        T* row = tab.data() + (size_t)(j - 1) * (nseg + 1) * 3;
        for (int c = 0; c <= nseg; ++c) {
            T* q = row + 3 * c;
            if (c < nseg && j <= deg[c]) {
                const double cc = (a[c] + b) / 2.0, e = (b - a[c]) / 2.0;
                if (j == 1) { q[0] = (T)(1.0 / e); q[1] = (T)(-cc / e); q[2] = T(0); }
                else { q[0] = (T)(2.0 / e); q[1] = (T)(-2.0 * cc / e); q[2] = T(-1); }
            } else {
                q[0] = T(0); q[1] = T(1); q[2] = T(0);  // copy through
            }
        }
    }
}

template <typename T>
struct Ring {
    T* b[3];
};

// Enqueue a segmented filter of `steps` steps on w columns (leading dimension
// ld) starting from x; returns the buffer holding the result.
template <typename T>
T* enqueue_filter(const DevLaplacian& L, Ring<T>& r, const T* x, int w, int ld, int steps,
                  const int32_t* rseg, const T* coef, int nseg, int64_t& spmm_calls,
                  hipStream_t s) {
    T* yprev = r.b[0];
    T* y = r.b[1];
    T* ynew = r.b[2];
    const size_t stride = (size_t)(nseg + 1) * 3;
    if constexpr (sizeof(T) == 4) spmm_seg_f32(L, x, y, w, ld, nullptr, rseg, coef, s);
    else spmm_seg_f64(L, x, y, w, ld, nullptr, rseg, coef, s);
    ++spmm_calls;
    const T* prev = x;
    for (int j = 2; j <= steps; ++j) {
        const T* cj = coef + (size_t)(j - 1) * stride;
        if constexpr (sizeof(T) == 4) spmm_seg_f32(L, y, ynew, w, ld, prev, rseg, cj, s);
        else spmm_seg_f64(L, y, ynew, w, ld, prev, rseg, cj, s);
        ++spmm_calls;
        T* t = yprev; yprev = y; y = ynew; ynew = t;
        prev = yprev;
    }
    return y;
}

// Segmented fused Rayleigh-Ritz of the first w columns of y (ld): ly = L y,
// G_c = y_c^T [y_c | ly_c], scaled generalized eig per segment, x/ax update and
// wanted-column norms.  Enqueue only (no synchronisation).
template <typename T>
void enqueue_rr(const DevLaplacian& L, T* y, T* ly, int w, int ld, int k, T* x, T* ax,
                const WorkList& gram, const WorkList& tiles, double* g, double* theta,
                double* cmat, int* status, double* norms2, double* gws, double* rws,
                int64_t& spmm_calls, hipStream_t s) {
    const int nseg = gram.w.nseg;
    if constexpr (sizeof(T) == 4) {
        spmm_f32(L, y, ly, w, ld, 1.f, 0.f, 0.f, nullptr, s);
        gram2_seg_f32(gram.w, y, ld, w, ly, ld, g, gws, s);
    } else {
        spmm_f64(L, y, ly, w, ld, 1.0, 0.0, 0.0, nullptr, s);
        gram2_seg_f64(gram.w, y, ld, w, ly, ld, g, gws, s);
    }
    ++spmm_calls;
    const int64_t gs = (int64_t)w * 2 * w;
    small_gen_eig_scaled_batched(nseg, w, g + w, 2 * w, gs, g, 2 * w, gs, theta, cmat, status, s);
    if constexpr (sizeof(T) == 4)
        rr_update_seg_f32(tiles.w, y, ly, ld, w, cmat, theta, k, x, ax, norms2, rws, s);
    else
        rr_update_seg_f64(tiles.w, y, ly, ld, w, cmat, theta, k, x, ax, norms2, rws, s);
}

}  // namespace

void fast_solve_segments(const DevLaplacian& L, const std::vector<int64_t>& row0,
                         const std::vector<int64_t>& rown, int n_components, uint64_t seed,
                         const FastOptions& opt, bool use_graphs, double* evals_host,
                         double* evecs_dev, std::vector<double>& resid, std::vector<int>& ok,
                         BatchStats& st, hipStream_t s) {
    const int nseg = (int)row0.size();
    const int64_t n = L.n;
    const int k = n_components + 1;
    st = BatchStats{};
    resid.assign(nseg, INFINITY);
    ok.assign(nseg, 0);
    if (nseg == 0) return;
    const int m = 16;
    for (int c = 0; c < nseg; ++c)
        SI_CHECK(rown[c] > 64 && row0[c] >= 0 && row0[c] + rown[c] <= n, kErrInvalidArgument,
                 "fast_solve_segments: segments must lie in the operator and hold > 64 rows");
    SI_CHECK(k <= 8 && k + 5 <= m, kErrUnsupported, "fast_solve_segments: n_components above 7");
    const int mf = std::min(m, std::max(8, k + 5));
    const int cap = opt.cheb_degree > 0 ? opt.cheb_degree : 24;
    const double b = 2.0;

    // row -> segment map (rows outside every segment use the frozen row nseg)
    DevBuf<int32_t> rseg(n);
    {
        std::vector<int32_t> h(n, nseg);
        for (int c = 0; c < nseg; ++c)
            for (int64_t r = 0; r < rown[c]; ++r) h[row0[c] + r] = c;
        to_device(rseg.get(), h.data(), h.size(), s);
    }
    WorkList gram, t32, t64;
    build_worklist(row0, rown, 1024, gram, s);
    build_worklist(row0, rown, 64, t32, s);   // FP32 update tiles (64 rows)
    build_worklist(row0, rown, 32, t64, s);   // FP64 update tiles (32 rows)

    const size_t nm = (size_t)n * m;
    float* x32 = ws_get<float>(kWsFastX32, nm);
    float* ax32 = ws_get<float>(kWsFastAx32, nm);
    Ring<float> r32{{ws_get<float>(kWsFastF32a, nm), ws_get<float>(kWsFastF32b, nm),
                     ws_get<float>(kWsFastF32c, nm)}};
    double* gws = ws_get<double>(kWsGram, (size_t)gram.w.nblk * 2 * m * m);
    double* rws = ws_get<double>(kWsBatchRr, (size_t)std::max(t32.w.nblk, t64.w.nblk) * 2 * k);
    const size_t gsz = (size_t)nseg * m * 2 * m, csz = (size_t)nseg * m * m;
    double* small = ws_get<double>(kWsBatchSmall, gsz + csz + (size_t)nseg * m + (size_t)nseg * 2 * k);
    double* g = small;
    double* cmat = g + gsz;
    double* theta = cmat + csz;
    double* norms2 = theta + (size_t)nseg * m;
    int* status = ws_get<int>(kWsStatus, nseg);
    const int max_steps = cap;
    float* coef32 = ws_get<float>(kWsBatchCoef, (size_t)max_steps * (nseg + 1) * 3);

    // start block: N(0,1) (index = original node id) with D^{1/2} 1 in column 0
    fill_normal_f32(x32, n, m, m, seed, kStreamLobpcgInit, s, L.perm);
    set_column_f32(x32, m, n, 0, L.sqrt_deg, 1.0, s);

    std::vector<double> a(nseg, 0.1), best(nseg, INFINITY), worst(nseg, INFINITY);
    std::vector<int> stall(nseg, 0), done(nseg, 0), deg(nseg, 0);
    std::vector<double> th((size_t)nseg * m), nrm((size_t)nseg * 2 * k);
    std::vector<int> hstat(nseg);
    std::vector<float> tab;
    std::map<int, hipGraphExec_t> graphs;
    const bool graphs_ok = use_graphs && !spmm_profiling();
    auto release_graphs = [&] {
        for (auto& kv : graphs) hipGraphExecDestroy(kv.second);
        graphs.clear();
    };
    // a round's host side: the coefficient table goes up through page-locked
    // memory without a wait, and the readback below is the round's only
    // synchronisation (SI_TRACE=1 reports the time spent waiting for the
    // device and the time spent in host decisions)
    using clk = std::chrono::steady_clock;
    double wait_ms = 0.0, host_ms = 0.0;
    clk::time_point th0 = clk::now();
    auto upload_async = [&](void* dst, const void* src, size_t bytes) {
        void* h = pinned_stage(bytes);
        std::memcpy(h, src, bytes);
        SI_HIP(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, s));
    };
    auto read_round = [&](int w) {
        const clk::time_point t0 = clk::now();
        host_ms += std::chrono::duration<double, std::milli>(t0 - th0).count();
        SI_HIP(hipMemcpyAsync(hstat.data(), status, sizeof(int) * nseg, hipMemcpyDeviceToHost, s));
        SI_HIP(hipMemcpyAsync(th.data(), theta, sizeof(double) * nseg * w, hipMemcpyDeviceToHost, s));
        SI_HIP(hipMemcpyAsync(nrm.data(), norms2, sizeof(double) * nseg * 2 * k,
                              hipMemcpyDeviceToHost, s));
        SI_HIP(hipStreamSynchronize(s));
        th0 = clk::now();
        wait_ms += std::chrono::duration<double, std::milli>(th0 - t0).count();
    };
    auto seg_worst = [&](int c) {
        double wv = 0.0;
        for (int j = 0; j < k; ++j) {
            double xn = std::sqrt(std::fmax(nrm[(size_t)c * 2 * k + k + j], 0.0));
            if (xn < 2.220446049250313e-16) xn = 2.220446049250313e-16;
            const double r = std::sqrt(std::fmax(nrm[(size_t)c * 2 * k + j], 0.0)) / xn;
            wv = std::isnan(r) ? INFINITY : std::max(wv, r);
        }
        return wv;
    };

    try {
        th0 = clk::now();
        double capture_ms = 0.0;
        // ---------------- FP32 rounds
        for (int round = 0; round < opt.max_outer; ++round) {
            int steps = 0;
            bool any = false;
            for (int c = 0; c < nseg; ++c) {
                deg[c] = done[c] ? 0 : capped_degree(a[c], b, cap);
                steps = std::max(steps, deg[c]);
                any = any || !done[c];
            }
            if (!any) break;
            coef_table<float>(a, deg, b, steps, tab);
            upload_async(coef32, tab.data(), sizeof(float) * tab.size());
            auto body = [&] {
                float* y = enqueue_filter<float>(L, r32, x32, m, m, steps, rseg, coef32, nseg,
                                                 st.spmm_calls, s);
                float* ly = (y == r32.b[0]) ? r32.b[1] : r32.b[0];
                enqueue_rr<float>(L, y, ly, m, m, k, x32, ax32, gram, t32, g, theta, cmat, status,
                                  norms2, gws, rws, st.spmm_calls, s);
            };
            if (graphs_ok) {
                auto it = graphs.find(steps);
                if (it == graphs.end()) {
                    const clk::time_point tc0 = clk::now();
                    hipGraph_t gr = nullptr;
                    hipGraphExec_t ex = nullptr;
                    const int64_t calls0 = st.spmm_calls;
                    SI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    try {
                        body();
                    } catch (...) {
                        hipStreamEndCapture(s, &gr);
                        if (gr) hipGraphDestroy(gr);
                        throw;
                    }
                    SI_HIP(hipStreamEndCapture(s, &gr));
                    SI_HIP(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
                    hipGraphDestroy(gr);
                    it = graphs.emplace(steps, ex).first;
                    ++st.graphs_captured;
                    st.spmm_calls = calls0;  // counted when launched
                    capture_ms += std::chrono::duration<double, std::milli>(clk::now() - tc0).count();
                }
                SI_HIP(hipGraphLaunch(it->second, s));
                ++st.graph_launches;
                st.spmm_calls += steps + 1;
            } else {
                body();
            }
            read_round(m);
            ++st.rounds;
            for (int c = 0; c < nseg; ++c) {
                if (done[c]) continue;
                if (hstat[c] != 0) { done[c] = 2; continue; }  // Y^T Y not positive definite
                const double wv = seg_worst(c);
                worst[c] = wv;
                if (wv <= opt.tol) { done[c] = 1; continue; }
                if (wv < 0.7 * best[c]) { best[c] = wv; stall[c] = 0; }
                else if (++stall[c] >= 3 && wv < 1e-4) { done[c] = 1; continue; }  // FP32 floor
                const double* tc = th.data() + (size_t)c * m;
                a[c] = std::min(std::max(tc[m - 1], tc[k - 1] * 1.5 + 1e-12), 0.6);
            }
        }
        release_graphs();

        // ---------------- FP64 finish on the lowest mf Ritz vectors
        double* x64 = ws_get<double>(kWsFastX64, nm);
        double* ax64 = ws_get<double>(kWsFastAx64, nm);
        Ring<double> r64{{ws_get<double>(kWsFastF64a, nm), ws_get<double>(kWsFastF64b, nm),
                          ws_get<double>(kWsFastF64c, nm)}};
        double* coef64 = ws_get<double>(kWsBatchCoef, (size_t)max_steps * (nseg + 1) * 3 * 2);
        convert_f32_to_f64(x32, r64.b[0], (int64_t)nm, s);
        enqueue_rr<double>(L, r64.b[0], r64.b[1], mf, m, k, x64, ax64, gram, t64, g, theta, cmat,
                           status, norms2, gws, rws, st.spmm_calls, s);
        read_round(mf);
        std::vector<int> fail(nseg, 0);
        for (int c = 0; c < nseg; ++c) {
            fail[c] = hstat[c] != 0;
            worst[c] = fail[c] ? INFINITY : seg_worst(c);
            done[c] = fail[c] || worst[c] <= opt.tol;
        }
        std::vector<double> a64(nseg, 0.1), thk((size_t)nseg * mf);
        std::copy(th.begin(), th.begin() + (size_t)nseg * mf, thk.begin());
        std::vector<double> tab64;
        for (int extra = 0; extra < 20; ++extra) {
            int steps = 0;
            for (int c = 0; c < nseg; ++c) {
                if (!done[c]) {
                    const double* tc = thk.data() + (size_t)c * mf;
                    a64[c] = std::min(std::max(tc[mf - 1], tc[k - 1] * 1.5 + 1e-12), 0.6);
                }
                deg[c] = done[c] ? 0 : capped_degree(a64[c], b, cap);
                steps = std::max(steps, deg[c]);
            }
            if (steps == 0) break;
            coef_table<double>(a64, deg, b, steps, tab64);
            upload_async(coef64, tab64.data(), sizeof(double) * tab64.size());
            double* y = enqueue_filter<double>(L, r64, x64, mf, m, steps, rseg, coef64, nseg,
                                               st.spmm_calls, s);
            double* ly = (y == r64.b[0]) ? r64.b[1] : r64.b[0];
            // the update rewrites x64/ax64 of every segment; converged segments
            // were copied through the filter, so they are re-projected unchanged
            enqueue_rr<double>(L, y, ly, mf, m, k, x64, ax64, gram, t64, g, theta, cmat, status,
                               norms2, gws, rws, st.spmm_calls, s);
            read_round(mf);
            ++st.rounds;
            for (int c = 0; c < nseg; ++c) {
                if (fail[c]) continue;
                if (hstat[c] != 0) { fail[c] = 1; done[c] = 1; worst[c] = INFINITY; continue; }
                std::copy(th.begin() + (size_t)c * mf, th.begin() + (size_t)(c + 1) * mf,
                          thk.begin() + (size_t)c * mf);
                worst[c] = seg_worst(c);
                if (worst[c] <= opt.tol) done[c] = 1;
            }
        }
        copy_cols_f64(x64, m, evecs_dev, k, n, k, s);
        SI_HIP(hipStreamSynchronize(s));
        if (trace_on())
            std::fprintf(stderr, "[batch] %d segments, %lld rounds: %.2f ms waiting for the device, "
                         "%.2f ms on the host between rounds (%.2f ms of it graph capture)\n",
                         nseg, (long long)st.rounds, wait_ms, host_ms, capture_ms);
        for (int c = 0; c < nseg; ++c) {
            for (int j = 0; j < k; ++j) evals_host[(size_t)c * k + j] = thk[(size_t)c * mf + j];
            resid[c] = worst[c];
            ok[c] = !fail[c] && worst[c] <= opt.tol;
        }
    } catch (...) {
        release_graphs();
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive) {
            hipGraph_t gr = nullptr;
            hipStreamEndCapture(s, &gr);
            if (gr) hipGraphDestroy(gr);
        }
        throw;
    }
}

}  // namespace si
