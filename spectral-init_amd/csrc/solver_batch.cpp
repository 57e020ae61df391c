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
//   * the decisions between rounds (convergence, stall and failure flags, each
//     segment's interval and degree, the next round's recurrence scalars) are
//     made on the device by the round planner (k_batch.hip), whose gate every
//     launch of a round reads, so several rounds are captured as one hipGraph
//     and replayed with no host work in between; the host only reads the gate
//     after each replay to see whether another is needed.
//
// The FP64 finish (Rayleigh-Ritz on the lowest mf Ritz vectors and the FP64
// residual against the exact Laplacian) is segmented and planned the same way.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "si_host.hpp"

namespace si {

namespace {

struct WorkList {
    SegWork w;
};

// Work items of at most `chunk` rows per segment, built on the host: rows
// (begin, end) pairs, the segment of each item, each segment's first item.
struct HostWork {
    std::vector<int64_t> rows;
    std::vector<int32_t> seg, segblk;
};
HostWork work_items(const std::vector<int64_t>& row0, const std::vector<int64_t>& rown,
                    int64_t chunk) {
    HostWork h;
    h.segblk.assign(row0.size() + 1, 0);
    for (size_t c = 0; c < row0.size(); ++c) {
        h.segblk[c] = (int32_t)h.seg.size();
        for (int64_t r = 0; r < rown[c]; r += chunk) {
            h.rows.push_back(row0[c] + r);
            h.rows.push_back(row0[c] + std::min(rown[c], r + chunk));
            h.seg.push_back((int32_t)c);
        }
    }
    h.segblk[row0.size()] = (int32_t)h.seg.size();
    return h;
}

template <typename T>
struct Ring {
    T* b[3];
};

// Enqueue a segmented filter of up to `steps` steps on w columns (leading
// dimension ld) starting from x; step j runs only while j <= *gate, and the
// ring puts the last step's result in r.b[0], which is returned.  Steps up to
// `always` run in every planned round (one workgroup per row block); later
// steps may idle, so they launch only the resident workgroups.
template <typename T>
T* enqueue_filter(const DevLaplacian& L, Ring<T>& r, const T* x, int w, int ld, int steps,
                  int always, const int32_t* rseg, const T* coef, int nseg, const int32_t* gate,
                  hipStream_t s, const T* lx = nullptr) {
    const size_t stride = (size_t)(nseg + 1) * 3;
    for (int j = 1; j <= steps; ++j) {
        const T* cj = coef + (size_t)(j - 1) * stride;
        if constexpr (sizeof(T) == 4) {
            if (j == 1 && lx)  // L x known from the previous update: no SpMM
                batch_axpby_ring_f32(L.n, w, ld, rseg, cj, lx, x, r.b, gate, s);
            else
                spmm_seg_ring_f32(L, x, r.b, w, ld, rseg, cj, gate, j, j > always, s);
        } else {
            spmm_seg_ring_f64(L, x, r.b, w, ld, rseg, cj, gate, j, j > always, s);
        }
    }
    return r.b[0];
}

// Segmented fused Rayleigh-Ritz of the first w columns of y (ld): ly = L y,
// G_c = y_c^T [y_c | ly_c], scaled generalized eig per segment, x/ax update and
// wanted-column norms.  Enqueue only (no synchronisation); with a gate, every
// launch returns at once when *gate == 0.
template <typename T>
void enqueue_rr(const DevLaplacian& L, T* y, T* ly, int w, int ld, int k, T* x, T* ax,
                const WorkList& gram, const WorkList& tiles, double* g, double* theta,
                double* cmat, int* status, double* norms2, double* gws, double* rws,
                const int32_t* gate, hipStream_t s) {
    const int nseg = gram.w.nseg;
    if constexpr (sizeof(T) == 4) {
        spmm_f32(L, y, ly, w, ld, 1.f, 0.f, 0.f, nullptr, s, gate, 1);
        gram2_seg_f32(gram.w, y, ld, w, ly, ld, g, gws, s, gate);
    } else {
        spmm_f64(L, y, ly, w, ld, 1.0, 0.0, 0.0, nullptr, s, gate, 1);
        gram2_seg_f64(gram.w, y, ld, w, ly, ld, g, gws, s, gate);
    }
    const int64_t gs = (int64_t)w * 2 * w;
    small_gen_eig_scaled_batched(nseg, w, g + w, 2 * w, gs, g, 2 * w, gs, theta, cmat, status, s,
                                 gate);
    if constexpr (sizeof(T) == 4)
        rr_update_seg_f32(tiles.w, y, ly, ld, w, cmat, theta, k, x, ax, norms2, rws, s, gate);
    else
        rr_update_seg_f64(tiles.w, y, ly, ld, w, cmat, theta, k, x, ax, norms2, rws, s, gate);
}

// Rounds captured per hipGraph: after the last useful round the gate closes
// and the remaining captured rounds cost only their empty launches.
constexpr int kRoundsPerGraph = 4;

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
    SI_CHECK(opt.max_outer >= 1, kErrInvalidArgument, "fast_solve_segments: max_outer >= 1");
    const int mf = std::min(m, std::max(8, k + 5));
    const int cap = opt.cheb_degree > 0 ? opt.cheb_degree : 24;
    // every round captures `steps` filter launches; the planner's degrees never exceed cap
    const int steps = std::max(cap, 2);
    // the planner's degree on [lo, 2] (as plan_degree in k_batch.hip)
    auto degree_at = [&](double lo) {
        const double x0 = (lo + 2.0) / (2.0 - lo);
        return std::max(2, std::min(steps, (int)std::floor(9.2 / std::acosh(x0))));
    };
    // the planner's smallest degree (its interval's lower end is at most 0.6):
    // the filter steps every planned round runs
    const int always = degree_at(0.6);
    // the first FP32 round: every segment starts at lo = 0.1, so every step
    // up to this degree runs
    const int first_steps = degree_at(0.1);
    const int max_extra64 = 20;

    // The work lists (Gram items of 1024 rows, FP32 / FP64 update tiles of 64 /
    // 32 rows) and the segments sorted by first row go up in one copy from
    // page-locked memory, into workspace (no allocation, no wait); the
    // row -> segment map is then built on the device.
    const HostWork hw[3] = {work_items(row0, rown, 1024), work_items(row0, rown, 64),
                            work_items(row0, rown, 32)};
    std::vector<int32_t> by_start(nseg);
    for (int c = 0; c < nseg; ++c) by_start[c] = c;
    std::sort(by_start.begin(), by_start.end(),
              [&](int32_t u, int32_t v) { return row0[u] < row0[v]; });
    size_t n64 = 2 * (size_t)nseg, n32 = (size_t)nseg;
    for (const HostWork& h : hw) {
        n64 += h.rows.size();
        n32 += h.seg.size() + h.segblk.size();
    }
    const size_t payload = 8 * n64 + 4 * n32;
    constexpr size_t kCtlBytes = 64;  // ctl readback area at the front of the stage
    char* stage = static_cast<char*>(pinned_stage(kCtlBytes + payload));
    char* dwork = ws_get<char>(kWsBatchWork, payload);
    WorkList wl[3];
    {
        int64_t* h64 = reinterpret_cast<int64_t*>(stage + kCtlBytes);
        int32_t* h32 = reinterpret_cast<int32_t*>(h64 + n64);
        const int64_t* d64 = reinterpret_cast<const int64_t*>(dwork);
        const int32_t* d32 = reinterpret_cast<const int32_t*>(d64 + n64);
        size_t o64 = 0, o32 = 0;
        auto put64 = [&](const int64_t* src, size_t cnt) {
            std::memcpy(h64 + o64, src, 8 * cnt);
            o64 += cnt;
            return d64 + o64 - cnt;
        };
        auto put32 = [&](const int32_t* src, size_t cnt) {
            std::memcpy(h32 + o32, src, 4 * cnt);
            o32 += cnt;
            return d32 + o32 - cnt;
        };
        for (int i = 0; i < 3; ++i) {
            wl[i].w.nseg = nseg;
            wl[i].w.nblk = (int64_t)hw[i].seg.size();
            wl[i].w.blk_rows = put64(hw[i].rows.data(), hw[i].rows.size());
            wl[i].w.blk_seg = put32(hw[i].seg.data(), hw[i].seg.size());
            wl[i].w.seg_blk = put32(hw[i].segblk.data(), hw[i].segblk.size());
        }
        std::vector<int64_t> starts(nseg), ends(nseg);
        for (int c = 0; c < nseg; ++c) {
            starts[c] = row0[by_start[c]];
            ends[c] = row0[by_start[c]] + rown[by_start[c]];
        }
        const int64_t* d_starts = put64(starts.data(), nseg);
        const int64_t* d_ends = put64(ends.data(), nseg);
        const int32_t* d_ids = put32(by_start.data(), nseg);
        SI_HIP(hipMemcpyAsync(dwork, stage + kCtlBytes, payload, hipMemcpyHostToDevice, s));
        batch_segment_map(ws_get<int32_t>(kWsBatchMap, (size_t)n), n, nseg, d_starts, d_ends,
                          d_ids, s);
    }
    const int32_t* rseg = ws_get<int32_t>(kWsBatchMap, (size_t)n);
    const WorkList& gram = wl[0];
    const WorkList& t32 = wl[1];
    const WorkList& t64 = wl[2];

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
    // coefficient tables: FP32 rounds use the first steps*(nseg+1)*3 floats,
    // FP64 rounds the whole buffer as doubles
    void* coef_buf = ws_get<double>(kWsBatchCoef, (size_t)steps * (nseg + 1) * 3);
    float* coef32 = static_cast<float*>(coef_buf);
    double* coef64 = static_cast<double*>(coef_buf);
    // planner state: a, best, worst [nseg] and thk [nseg x mf] doubles, then
    // stall, done, deg [nseg] and ctl [4] ints
    const size_t plan_doubles = (size_t)nseg * (3 + mf);
    const size_t plan_ints = (size_t)nseg * 3 + 4;
    double* pd = ws_get<double>(kWsBatchPlan, plan_doubles + (plan_ints + 1) / 2);
    int32_t* pi = reinterpret_cast<int32_t*>(pd + plan_doubles);
    BatchPlan P;
    P.nseg = nseg;
    P.k = k;
    P.cap = cap;
    P.steps = steps;
    P.tol = opt.tol;
    P.b = 2.0;
    P.a = pd;
    P.best = pd + nseg;
    P.worst = pd + 2 * (size_t)nseg;
    P.stall = pi;
    P.done = pi + nseg;
    P.deg = pi + 2 * (size_t)nseg;
    P.ctl = pi + 3 * (size_t)nseg;
    P.status = status;
    P.theta = theta;
    P.norms2 = norms2;

    // start block: N(0,1) (index = original node id) with D^{1/2} 1 in column 0
    fill_normal_f32(x32, n, m, m, seed, kStreamLobpcgInit, s, L.perm);
    set_column_f32(x32, m, n, 0, L.sqrt_deg, 1.0, s);

    const bool graphs_ok = use_graphs && !spmm_profiling();
    // the front of the page-locked stage (its work-list part is read by the
    // copy above, which the stream completes before any readback)
    int32_t* ctl_host = reinterpret_cast<int32_t*>(stage);
    using clk = std::chrono::steady_clock;
    double wait_ms = 0.0, capture_ms = 0.0;
    int host_reads = 0;
    auto read_ctl = [&] {
        const clk::time_point t0 = clk::now();
        SI_HIP(hipMemcpyAsync(ctl_host, P.ctl, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        SI_HIP(hipStreamSynchronize(s));
        wait_ms += std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        ++host_reads;
    };
    hipGraphExec_t exec = nullptr;
    auto release = [&] {
        if (exec) hipGraphExecDestroy(exec);
        exec = nullptr;
    };
    // Captures `count` rounds (the first one by `first`) as one graph.
    auto capture = [&](hipGraphExec_t& ex, const auto& first, const auto& round, int count) {
        const clk::time_point tc0 = clk::now();
        hipGraph_t gr = nullptr;
        SI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            first();
            for (int r = 1; r < count; ++r) round();
        } catch (...) {
            hipStreamEndCapture(s, &gr);
            if (gr) hipGraphDestroy(gr);
            throw;
        }
        SI_HIP(hipStreamEndCapture(s, &gr));
        const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        hipGraphDestroy(gr);
        SI_HIP(e);
        ++st.graphs_captured;
        capture_ms += std::chrono::duration<double, std::milli>(clk::now() - tc0).count();
    };
    // Replays rounds until the planner closes the gate: kRoundsPerGraph rounds
    // per graph launch (the first launch starts with `first`, later ones with
    // `round`), or one eagerly enqueued round per host read.  The planner
    // closes the gate after max_rounds planned rounds, so the loop ends; the
    // host bound is a second guard.
    auto run_rounds = [&](const auto& first, const auto& round, int max_rounds) {
        read_ctl();
        if (ctl_host[0] == 0) return;
        const int per = graphs_ok ? kRoundsPerGraph : 1;
        const int max_launches = max_rounds / per + 2;
        for (int launch = 0; launch < max_launches; ++launch) {
            if (graphs_ok) {
                if (launch == 0) {
                    capture(exec, first, round, per);
                } else if (launch == 1) {
                    release();
                    capture(exec, round, round, per);
                }
                SI_HIP(hipGraphLaunch(exec, s));
                ++st.graph_launches;
            } else if (launch == 0) {
                first();
            } else {
                round();
            }
            read_ctl();
            if (ctl_host[0] == 0) break;
        }
        release();
        SI_CHECK(ctl_host[0] == 0, kErrInternal, "batched solve: the round planner did not stop");
    };

    try {
        // ---------------- FP32 rounds
        P.m = m;
        P.thk = nullptr;
        P.max_rounds = opt.max_outer;
        batch_plan_init(P, s);
        {
            BatchPlan p1 = P;  // the first round runs first_steps steps
            p1.steps = first_steps;
            batch_plan_f32(p1, 0, coef32, s);
        }
        // the first round filters the start block; later ones start from the
        // L X that the previous round's update left in ax32
        auto round32 = [&](int nsteps, int full, bool lx) {
            float* y = enqueue_filter<float>(L, r32, x32, m, m, nsteps, full, rseg, coef32, nseg,
                                             P.ctl, s, lx ? ax32 : nullptr);
            float* ly = r32.b[1];
            enqueue_rr<float>(L, y, ly, m, m, k, x32, ax32, gram, t32, g, theta, cmat, status,
                              norms2, gws, rws, P.ctl, s);
            batch_plan_f32(P, 1, coef32, s);  // plans the next round over `steps` steps
        };
        run_rounds([&] { round32(first_steps, first_steps, false); },
                   [&] { round32(steps, always, true); }, P.max_rounds);
        st.rounds += ctl_host[1];
        st.spmm_calls += ctl_host[2];

        // ---------------- FP64 finish on the lowest mf Ritz vectors
        double* x64 = ws_get<double>(kWsFastX64, nm);
        double* ax64 = ws_get<double>(kWsFastAx64, nm);
        Ring<double> r64{{ws_get<double>(kWsFastF64a, nm), ws_get<double>(kWsFastF64b, nm),
                          ws_get<double>(kWsFastF64c, nm)}};
        convert_f32_to_f64(x32, r64.b[0], (int64_t)nm, s);
        enqueue_rr<double>(L, r64.b[0], r64.b[1], mf, m, k, x64, ax64, gram, t64, g, theta, cmat,
                           status, norms2, gws, rws, nullptr, s);
        st.spmm_calls += 1;
        P.m = mf;
        P.thk = pd + 3 * (size_t)nseg;
        P.max_rounds = max_extra64;
        batch_plan_init(P, s);
        // the update rewrites x64/ax64 of every segment; converged segments
        // were copied through the filter, so they are re-projected unchanged
        batch_plan_f64(P, 2, coef64, s);
        auto round64 = [&] {
            double* y = enqueue_filter<double>(L, r64, x64, mf, m, steps, always, rseg, coef64,
                                               nseg, P.ctl, s);
            double* ly = r64.b[1];
            enqueue_rr<double>(L, y, ly, mf, m, k, x64, ax64, gram, t64, g, theta, cmat, status,
                               norms2, gws, rws, P.ctl, s);
            batch_plan_f64(P, 1, coef64, s);
        };
        run_rounds(round64, round64, P.max_rounds);
        st.rounds += ctl_host[1];
        st.spmm_calls += ctl_host[2];
        // only the segments' rows: rows of other components belong to other
        // solvers (the dense EVD of the small ones may be writing them)
        batch_copy_rows(t64.w, x64, m, evecs_dev, k, k, s);

        std::vector<double> thk((size_t)nseg * mf), worst(nseg);
        std::vector<int32_t> done(nseg);
        to_host(thk.data(), P.thk, thk.size(), s);
        to_host(worst.data(), P.worst, worst.size(), s);
        to_host(done.data(), P.done, done.size(), s);  // synchronises the stream
        if (trace_on())
            std::fprintf(stderr, "[batch] %d segments, %d rounds in %d graph launches (%d host "
                         "reads): %.2f ms waiting for the device, %.2f ms graph capture\n",
                         nseg, st.rounds, st.graph_launches, host_reads, wait_ms, capture_ms);
        for (int c = 0; c < nseg; ++c) {
            for (int j = 0; j < k; ++j) evals_host[(size_t)c * k + j] = thk[(size_t)c * mf + j];
            const bool fail = done[c] == 2;
            resid[c] = fail ? INFINITY : worst[c];
            ok[c] = !fail && worst[c] <= opt.tol;
        }
    } catch (...) {
        release();
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
