// solver_common.cpp -- FP64 building blocks shared by the solvers: residual
// norms, Rayleigh-Ritz refinement and the Chebyshev filter.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <thread>

#include "si_host.hpp"

namespace si {

namespace {
struct Workspace {
    std::vector<DevBuf<char>> slots;
    Workspace() : slots(kWsNumSlots) {}
};
thread_local Workspace* g_ws = nullptr;
// page-locked host staging for the solvers' small per-round downloads (a copy
// into pageable memory goes through the runtime's own staging blits)
struct PinnedStage {
    void* p = nullptr;
    size_t n = 0;
};
thread_local PinnedStage* g_pin = nullptr;
// page-locked chunks of the staged uploads, each with the event of its last DMA
struct UploadRing {
    static constexpr int kSlots = 4;
    static constexpr size_t kChunk = (size_t)16 << 20;
    char* slot[kSlots] = {};
    hipEvent_t ev[kSlots] = {};
    bool pending[kSlots] = {};
    bool ok = false;
    UploadRing() {
        for (int i = 0; i < kSlots; ++i) {
            if (hipHostMalloc(reinterpret_cast<void**>(&slot[i]), kChunk, hipHostMallocDefault) !=
                    hipSuccess ||
                hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess)
                return;
        }
        ok = true;
    }
    ~UploadRing() {
        for (int i = 0; i < kSlots; ++i) {
            if (ev[i]) {
                if (pending[i]) (void)hipEventSynchronize(ev[i]);
                (void)hipEventDestroy(ev[i]);
            }
            if (slot[i]) (void)hipHostFree(slot[i]);
        }
    }
};
thread_local UploadRing* g_ring = nullptr;
}  // namespace

void upload_to_device(void* dst, const void* src, size_t bytes, hipStream_t s) {
    constexpr size_t kStagedMin = (size_t)8 << 20;
    if (bytes < kStagedMin) {
        SI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
        return;
    }
    if (!g_ring) g_ring = new UploadRing();
    if (!g_ring->ok) {  // no page-locked memory: the runtime's pageable path
        SI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
        return;
    }
    UploadRing& r = *g_ring;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nth = (int)std::min(4u, hw);
    const char* from = static_cast<const char*>(src);
    char* to = static_cast<char*>(dst);
    for (size_t off = 0, k = 0; off < bytes; off += UploadRing::kChunk, ++k) {
        const size_t len = std::min(UploadRing::kChunk, bytes - off);
        const int sl = (int)(k % UploadRing::kSlots);
        if (r.pending[sl]) {
            SI_HIP(hipEventSynchronize(r.ev[sl]));
            r.pending[sl] = false;
        }
        // the chunk's host copy, split over nth threads (this one included)
        char* pin = r.slot[sl];
        auto part = [&](int t) {
            const size_t a = len * t / nth, b = len * (t + 1) / nth;
            std::memcpy(pin + a, from + off + a, b - a);
        };
        std::vector<std::thread> pool;
        try {
            for (int t = 1; t < nth; ++t) pool.emplace_back(part, t);
        } catch (...) {
            for (auto& th : pool) th.join();
            throw;
        }
        part(0);
        for (auto& th : pool) th.join();
        SI_HIP(hipMemcpyAsync(to + off, pin, len, hipMemcpyHostToDevice, s));
        SI_HIP(hipEventRecord(r.ev[sl], s));
        r.pending[sl] = true;
    }
}

void* pinned_stage(size_t bytes) {
    if (!g_pin) g_pin = new PinnedStage();
    if (g_pin->n < bytes) {
        if (g_pin->p) SI_HIP(hipHostFree(g_pin->p));
        g_pin->p = nullptr;
        g_pin->n = 0;
        SI_HIP(hipHostMalloc(&g_pin->p, bytes, hipHostMallocDefault));
        g_pin->n = bytes;
    }
    return g_pin->p;
}

void stream_pool_setup() {
    constexpr int kMaxDevices = 64;
    static std::once_flag once[kMaxDevices];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return;
    std::call_once(once[dev], [dev] {
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess || pool == nullptr) return;
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    });
}

hipStream_t aux_stream(Context& cx) {
    if (!cx.aux) SI_HIP(hipStreamCreateWithFlags(&cx.aux, hipStreamNonBlocking));
    return cx.aux;
}

void* workspace_get(int slot, size_t bytes) {
    if (!g_ws) g_ws = new Workspace();
    DevBuf<char>& b = g_ws->slots[slot];
    if (b.p == nullptr || b.n < bytes) {
        // grow by 25 % beyond the request so nearby sizes reuse the block
        b.alloc(bytes + bytes / 4);
    }
    return b.p;
}

void download_to_host(void* dst, const void* src, size_t bytes, hipStream_t s) {
    constexpr size_t kStagedMin = (size_t)8 << 20;
    if (bytes >= kStagedMin && !g_ring) g_ring = new UploadRing();
    if (bytes < kStagedMin || !g_ring->ok) {
        SI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        SI_HIP(hipStreamSynchronize(s));
        return;
    }
    UploadRing& r = *g_ring;
    const size_t chunk = UploadRing::kChunk;
    const size_t nchunks = (bytes + chunk - 1) / chunk;
    const char* from = static_cast<const char*>(src);
    char* to = static_cast<char*>(dst);
    auto enqueue = [&](size_t k) {
        const int sl = (int)(k % UploadRing::kSlots);
        if (r.pending[sl]) {
            SI_HIP(hipEventSynchronize(r.ev[sl]));
            r.pending[sl] = false;
        }
        const size_t off = k * chunk, len = std::min(chunk, bytes - off);
        SI_HIP(hipMemcpyAsync(r.slot[sl], from + off, len, hipMemcpyDeviceToHost, s));
        SI_HIP(hipEventRecord(r.ev[sl], s));
        r.pending[sl] = true;
    };
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nth = (int)std::min(4u, hw);
    for (size_t k = 0; k < std::min<size_t>(nchunks, UploadRing::kSlots); ++k) enqueue(k);
    for (size_t k = 0; k < nchunks; ++k) {
        const int sl = (int)(k % UploadRing::kSlots);
        SI_HIP(hipEventSynchronize(r.ev[sl]));
        r.pending[sl] = false;
        const size_t off = k * chunk, len = std::min(chunk, bytes - off);
        const char* pin = r.slot[sl];
        auto part = [&](int t) {
            const size_t a = len * t / nth, b = len * (t + 1) / nth;
            std::memcpy(to + off + a, pin + a, b - a);
        };
        std::vector<std::thread> pool;
        try {
            for (int t = 1; t < nth; ++t) pool.emplace_back(part, t);
        } catch (...) {
            for (auto& th : pool) th.join();
            throw;
        }
        part(0);
        for (auto& th : pool) th.join();
        if (k + UploadRing::kSlots < nchunks) enqueue(k + UploadRing::kSlots);
    }
}

void workspace_release() {
    delete g_ws;
    g_ws = nullptr;
    delete g_ring;
    g_ring = nullptr;
    if (g_pin) {
        if (g_pin->p) (void)hipHostFree(g_pin->p);
        delete g_pin;
        g_pin = nullptr;
    }
}

// eigenpair_residual (metrics.rs:57-68) of column c from the FP64 Gram
// diagonals rr = ||L v - lambda v||^2 and vv = ||v||^2: the norm is floored at
// 1e-300 as there, and a NaN in either sum (a NaN entry of v, L v or lambda)
// gives a NaN residual, which max_eigenpair_residual (metrics.rs:73-92)
// propagates.  rr and vv are sums of squares, never negative unless NaN.
double residual_from_sums(double rr, double vv) {
    if (std::isnan(rr) || std::isnan(vv)) return NAN;
    double vn = std::sqrt(vv);
    if (vn < 1e-300) vn = 1e-300;
    return std::sqrt(rr) / vn;
}

// ||L v_j - lambda_j v_j|| / max(||v_j||, 1e-300): one block SpMM, one residual
// block, two FP64 Grams (metrics.rs:57-68, batched over the k columns as
// per_vector_residuals does, lobpcg.rs:193-223).
void fp64_residuals(const DevLaplacian& L, const double* v, int k, const double* evals_host,
                    double* res_host, hipStream_t s) {
    const int64_t n = L.n;
    double* ax = ws_get<double>(kWsResAx, (size_t)n * k);
    double* r = ws_get<double>(kWsResR, (size_t)n * k);
    double* lam = ws_get<double>(kWsSmall0, k);
    double* g = ws_get<double>(kWsSmall1, (size_t)k * k);
    double* gx = ws_get<double>(kWsSmall2, (size_t)k * k);
    double* ws = ws_get<double>(kWsGram, gram_workspace_doubles(n, k, k));
    spmm_f64(L, v, ax, k, k, 1.0, 0.0, 0.0, nullptr, s);
    to_device(lam, evals_host, k, s);
    residual_f64(n, ax, v, k, k, lam, r, k, s);
    gram_f64(n, r, k, k, r, k, k, g, ws, s);
    gram_f64(n, v, k, k, v, k, k, gx, ws, s);
    std::vector<double> hg((size_t)k * k), hx((size_t)k * k);
    to_host(hg.data(), g, hg.size(), s);
    to_host(hx.data(), gx, hx.size(), s);
    for (int c = 0; c < k; ++c) res_host[c] = residual_from_sums(hg[c * k + c], hx[c * k + c]);
}

// rayleigh_ritz_refine (lobpcg.rs:158-187): AX = L X, G = X^T AX, eigh (lower
// triangle), X <- X V.
bool rayleigh_ritz_f64(const DevLaplacian& L, double* v, int k, double* evals_host, hipStream_t s) {
    const int64_t n = L.n;
    double* ax = ws_get<double>(kWsRrAx, (size_t)n * k);
    double* g = ws_get<double>(kWsSmall3, (size_t)k * k);
    double* w = ws_get<double>(kWsSmall4, k);
    double* vv = ws_get<double>(kWsSmall5, (size_t)k * k);
    double* ws = ws_get<double>(kWsGram, gram_workspace_doubles(n, k, k));
    spmm_f64(L, v, ax, k, k, 1.0, 0.0, 0.0, nullptr, s);
    gram_f64(n, v, k, k, ax, k, k, g, ws, s);
    std::vector<double> hg((size_t)k * k);
    to_host(hg.data(), g, hg.size(), s);
    for (double x : hg)
        if (!std::isfinite(x)) return false;
    small_sym_eig(k, g, k, w, vv, s);
    block_mul_f64(n, v, k, k, vv, k, nullptr, 0, 0, nullptr, v, k, s);
    to_host(evals_host, w, k, s);
    return true;
}

// chebyshev_precond (lobpcg.rs:113-150) through the fused SpMM epilogue:
//   y1 = (L r - c r) / e,   y_{j+1} = scale L y_j - shift y_j - y_{j-1}.
void chebyshev_filter_f64(const DevLaplacian& L, const double* r, int k, double a, double b,
                          int degree, double* out, hipStream_t s) {
    const int64_t n = L.n;
    const size_t cnt = (size_t)n * k;
    const double c = (a + b) / 2.0, e = (b - a) / 2.0;
    const double scale = 2.0 / e, shift = scale * c;
    if (degree <= 0) {
        SI_HIP(hipMemcpyAsync(out, r, sizeof(double) * cnt, hipMemcpyDeviceToDevice, s));
        return;
    }
    double* b0 = ws_get<double>(kWsFastF64a, cnt);
    double* b1 = ws_get<double>(kWsFastF64b, cnt);
    double* b2 = ws_get<double>(kWsFastF64c, cnt);
    SI_HIP(hipMemcpyAsync(b0, r, sizeof(double) * cnt, hipMemcpyDeviceToDevice, s));
    spmm_f64(L, b0, b1, k, k, 1.0 / e, -c / e, 0.0, nullptr, s);
    double* yprev = b0;
    double* y = b1;
    double* ynew = b2;
    for (int d = 2; d <= degree; ++d) {
        spmm_f64(L, y, ynew, k, k, scale, -shift, -1.0, yprev, s);
        double* t = yprev; yprev = y; y = ynew; ynew = t;
    }
    SI_HIP(hipMemcpyAsync(out, y, sizeof(double) * cnt, hipMemcpyDeviceToDevice, s));
    SI_HIP(hipStreamSynchronize(s));
}

}  // namespace si
