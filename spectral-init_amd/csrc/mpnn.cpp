// mpnn.cpp -- the learned warm-start network on the device: node features,
// forward pass, loss gradient, backward pass, Adam step and the
// residual-judged warm-start block, plus their C ABI.
//
// The reference starts LOBPCG from N(0,1) plus the sqrt-degree hint
// (src/solvers/lobpcg.rs:243-252); here a 3-layer message-passing network
// (hidden 64, tanh; node features: degree, local normalised-weight statistics
// and 8 random-walk probes) maps each node to 8 outputs whose span should
// contain the two non-trivial eigenvectors (a subspace-coverage loss).  The
// warm start evaluates it on three independent probe sets and takes the
// Rayleigh-Ritz vectors of [D^{1/2} 1 | the 24 outputs].  Dense layers run on
// the bf16 matrix cores (k_mpnn.hip), the neighbour aggregation
// A_hat H = (I - L) H on the FP32 SpMM, the weight gradients on the FP64 Gram
// kernel.  The block it proposes is used only if its Rayleigh-Ritz residual
// beats a random start's.
#include <algorithm>
#include <cmath>
#include <vector>

#include "si_abi.hpp"
#include "../../include/spectral_mi355x_warmstart.h"

namespace si {

namespace {

constexpr int kHid = 64, kNf = 16, kNout = kMpnnOut;
// independent probe sets the warm start evaluates the network on
constexpr int kProbeSets = 3;
// ridge of the coverage loss: G + kRidge tr(G) / kNout I
constexpr double kRidge = 1e-3;
// walk lengths at which the 8 random-walk probes are recorded (multi-scale)
constexpr int kProbeSteps[8] = {1, 2, 3, 4, 6, 8, 12, 16};

struct Layout {
    int64_t off[10];
    int64_t total;
};

Layout layout() {
    Layout l{};
    int64_t o = 0;
    int din = kNf;
    for (int j = 0; j < 3; ++j) {
        l.off[3 * j] = o; o += (int64_t)din * kHid;
        l.off[3 * j + 1] = o; o += (int64_t)din * kHid;
        l.off[3 * j + 2] = o; o += kHid;
        din = kHid;
    }
    l.off[9] = o;
    o += kHid * kNout;
    l.total = o;
    return l;
}

}  // namespace

struct Mpnn {
    Layout lay = layout();
    DevBuf<float> p, m, v;   // master parameters, Adam moments (FP32)
    DevBuf<double> grad;     // gradient (FP64)
    DevBuf<short> w[3][4];   // layer l: Ws, Wn (bf16 [din x 64]), WsT, WnT (bf16 [64 x din])
    DevBuf<double> wo, wot;  // output weights [64 x 8] and [8 x 64] for the block products
    int step = 0;
    bool packed = false;
    int64_t cap = 0;
    DevBuf<float> h[3], av[3], x, dx, dz0, dz1, u, t, ones;
    DevBuf<double> xd;
};

namespace {

void ensure(Mpnn& M, int64_t n, hipStream_t s) {
    if (n <= M.cap) return;
    for (int l = 0; l < 3; ++l) { M.h[l].alloc((size_t)n * kHid); M.av[l].alloc((size_t)n * kHid); }
    M.x.alloc((size_t)n * kNout);
    M.dx.alloc((size_t)n * kNout);
    M.dz0.alloc((size_t)n * kHid);
    M.dz1.alloc((size_t)n * kHid);
    M.u.alloc((size_t)n * kHid);
    M.t.alloc((size_t)n * kHid);
    M.xd.alloc((size_t)n * kNout);
    M.ones.alloc(n);
    std::vector<float> one(n, 1.f);
    to_device(M.ones.get(), one.data(), one.size(), s);
    M.cap = n;
}

void pack(Mpnn& M, hipStream_t s) {
    if (M.packed) return;
    int din = kNf;
    for (int l = 0; l < 3; ++l) {
        for (int j = 0; j < 2; ++j) {
            const float* src = M.p.get() + M.lay.off[3 * l + j];
            if (!M.w[l][j].get()) M.w[l][j].alloc((size_t)din * kHid);
            if (!M.w[l][j + 2].get()) M.w[l][j + 2].alloc((size_t)din * kHid);
            mpnn_pack(src, din, kHid, 0, M.w[l][j], s);
            mpnn_pack(src, din, kHid, 1, M.w[l][j + 2], s);
        }
        din = kHid;
    }
    std::vector<float> hwo(kHid * kNout);
    to_host(hwo.data(), M.p.get() + M.lay.off[9], hwo.size(), s);
    std::vector<double> a(kHid * kNout), b(kHid * kNout);
    for (int k = 0; k < kHid; ++k)
        for (int c = 0; c < kNout; ++c) {
            a[k * kNout + c] = hwo[k * kNout + c];
            b[c * kHid + k] = hwo[k * kNout + c];
        }
    if (!M.wo.get()) { M.wo.alloc(a.size()); M.wot.alloc(b.size()); }
    to_device(M.wo.get(), a.data(), a.size(), s);
    to_device(M.wot.get(), b.data(), b.size(), s);
    M.packed = true;
}

// A_hat x = (I - L) x through the FP32 SpMM epilogue (s_l = -1, s_x = 1)
void ahat(const DevLaplacian& L, const float* x, float* y, int k, hipStream_t s) {
    spmm_f32(L, x, y, k, k, -1.f, 1.f, 0.f, nullptr, s);
}

}  // namespace

void mpnn_init(Mpnn& M, uint64_t seed, hipStream_t s) {
    const int64_t P = M.lay.total;
    std::vector<float> h(P, 0.f);
    // Glorot-uniform weights from the shared Philox stream; zero biases
    int din = kNf;
    for (int l = 0; l <= 3; ++l) {
        const int fin = l < 3 ? din : kHid, fout = l < 3 ? kHid : kNout;
        const double lim = std::sqrt(6.0 / (fin + fout));
        const int blocks = l < 3 ? 2 : 1;
        for (int j = 0; j < blocks; ++j) {
            const int64_t o = l < 3 ? M.lay.off[3 * l + j] : M.lay.off[9];
            for (int64_t e = 0; e < (int64_t)fin * fout; ++e) {
                const double u = 0.5 * (1.0 + std::erf(normal_at(seed, kStreamNetInit, (uint64_t)(o + e)) / std::sqrt(2.0)));
                h[o + e] = (float)((2.0 * u - 1.0) * lim);
            }
        }
        din = kHid;
    }
    M.p.alloc(P); M.m.alloc(P); M.v.alloc(P); M.grad.alloc(P);
    to_device(M.p.get(), h.data(), P, s);
    std::vector<float> z(P, 0.f);
    to_device(M.m.get(), z.data(), P, s);
    to_device(M.v.get(), z.data(), P, s);
    M.step = 0;
    M.packed = false;
}

// Node features [n x 16] in the Laplacian's row order (see oracle_mpnn.c) of
// `sets` probe seeds seed, seed + 1, ... at once, into f = [sets][n x 16].  The
// random-walk probes of all sets advance together, one `sets`-column FP64 SpMM
// per step; each column's arithmetic is that of a walk of its own, so set j's
// features are those of a single-set call with seed + j.
void mpnn_features_sets(const DevLaplacian& L, uint64_t seed, int sets, float* f, hipStream_t s) {
    const int64_t n = L.n;
    const size_t nf = (size_t)n * 16;
    std::vector<double> sq(n);
    to_host(sq.data(), L.sqrt_deg, n, s);
    double sqmax = 0.0, sqn2 = 0.0;
    for (double v : sq) { sqmax = std::max(sqmax, v); sqn2 += v * v; }
    mpnn_rowfeat(L, sqmax > 0.0 ? 1.0 / sqmax : 0.0, f, s);
    for (int j = 1; j < sets; ++j)  // the degree and weight columns do not depend on the seed
        SI_HIP(hipMemcpyAsync(f + j * nf, f, sizeof(float) * nf, hipMemcpyDeviceToDevice, s));
    // the walk runs in FP64: A_hat never damps the trivial direction, so FP32
    // rounding noise along it would grow relative to the decaying probe
    DevBuf<double> p((size_t)n * sets), q((size_t)n * sets), one(n), g((size_t)sets * sets),
        ws(gram_workspace_doubles(n, sets, sets));
    DevBuf<float> pf(n);
    for (int j = 0; j < sets; ++j) {
        fill_normal_f64(one, n, 1, 1, seed + (uint64_t)j, kStreamProbe, s, L.perm);
        gram_f64(n, one, 1, 1, L.sqrt_deg, 1, 1, g, ws, s);
        double dot = 0.0;
        to_host(&dot, g.get(), 1, s);
        mpnn_deflate(n, one, L.sqrt_deg, sqn2 > 0.0 ? dot / sqn2 : 0.0, pf, s);
        copy_cols_f64(one, 1, p.get() + j, sets, n, 1, s);
    }
    double* cur = p.get();
    double* nxt = q.get();
    int step = 0;
    std::vector<double> hg((size_t)sets * sets);
    for (int t = 0; t < 8; ++t) {
        for (; step < kProbeSteps[t]; ++step) {
            spmm_f64(L, cur, nxt, sets, sets, -1.0, 1.0, 0.0, nullptr, s);  // A_hat = I - L
            std::swap(cur, nxt);
        }
        gram_f64(n, cur, sets, sets, cur, sets, sets, g, ws, s);
        to_host(hg.data(), g.get(), hg.size(), s);
        for (int j = 0; j < sets; ++j) {
            const double rms = std::sqrt(hg[(size_t)j * sets + j] / (double)n);
            copy_cols_f64(cur + j, sets, one, 1, n, 1, s);
            convert_f64_to_f32(one, pf, n, s);
            mpnn_setcol(n, 6 + t, pf, rms > 0.0 ? (float)(1.0 / rms) : 0.f, f + j * nf, s);
        }
    }
    SI_HIP(hipStreamSynchronize(s));
}

void mpnn_features(const DevLaplacian& L, uint64_t seed, float* f, hipStream_t s) {
    mpnn_features_sets(L, seed, 1, f, s);
}

// X = net(F) [n x 8], activations kept for the backward pass
void mpnn_forward(Mpnn& M, const DevLaplacian& L, const float* f, hipStream_t s) {
    const int64_t n = L.n;
    ensure(M, n, s);
    pack(M, s);
    const float* hin = f;
    int din = kNf;
    for (int l = 0; l < 3; ++l) {
        ahat(L, hin, M.av[l], din, s);
        mpnn_gemm(n, din, kHid, hin, M.w[l][0], M.av[l], M.w[l][1], M.p.get() + M.lay.off[3 * l + 2],
                  nullptr, 1, nullptr, M.h[l], s);
        hin = M.h[l];
        din = kHid;
    }
    block_mul_f32(n, M.h[2], kHid, kHid, M.wo, kNout, nullptr, 0, 0, nullptr, M.x, kNout, s);
}

// Cholesky solve of the SPD system a [m x m] c = b [m x r], b overwritten by c
bool chol_solve_small(int m, double* a, double* b, int r) {
    for (int j = 0; j < m; ++j) {
        double d = a[j * m + j];
        for (int k = 0; k < j; ++k) d -= a[j * m + k] * a[j * m + k];
        if (!(d > 0.0)) return false;
        d = std::sqrt(d);
        a[j * m + j] = d;
        for (int i = j + 1; i < m; ++i) {
            double e = a[i * m + j];
            for (int k = 0; k < j; ++k) e -= a[i * m + k] * a[j * m + k];
            a[i * m + j] = e / d;
        }
    }
    for (int c = 0; c < r; ++c) {
        for (int i = 0; i < m; ++i) {
            double e = b[i * r + c];
            for (int k = 0; k < i; ++k) e -= a[i * m + k] * b[k * r + c];
            b[i * r + c] = e / a[i * m + i];
        }
        for (int i = m - 1; i >= 0; --i) {
            double e = b[i * r + c];
            for (int k = i + 1; k < m; ++k) e -= a[k * m + i] * b[k * r + c];
            b[i * r + c] = e / a[i * m + i];
        }
    }
    return true;
}

// Coverage loss of the target V [n x 2] (FP64, Laplacian row order) by
// span(X), X = the network's 8 outputs, and its gradient in M.dx:
//   G = X^T X, B = X^T V, delta = kRidge tr(G) / 8, C = (G + delta I)^-1 B,
//   loss = tr(V^T V) - tr(B^T C),
//   dX = X (2 C C^T + (2 kRidge / 8) tr(C^T C) I) - 2 V C^T.
double mpnn_loss(Mpnn& M, int64_t n, const double* vt, hipStream_t s) {
    constexpr int P = kNout;
    convert_f32_to_f64(M.x, M.xd, n * P, s);
    double* g = ws_get<double>(kWsSmall6, P * P + 2 * P + 4);
    double* ws = ws_get<double>(kWsGram, gram_workspace_doubles(n, P, P));
    gram_f64(n, M.xd, P, P, M.xd, P, P, g, ws, s);
    gram_f64(n, M.xd, P, P, vt, 2, 2, g + P * P, ws, s);
    gram_f64(n, vt, 2, 2, vt, 2, 2, g + P * P + 2 * P, ws, s);
    std::vector<double> h(P * P + 2 * P + 4);
    to_host(h.data(), g, h.size(), s);
    const double* G = h.data();
    const double* B = h.data() + P * P;
    const double vv = h[P * P + 2 * P] + h[P * P + 2 * P + 3];
    double tr = 0.0;
    for (int p = 0; p < P; ++p) tr += G[p * P + p];
    std::vector<double> gd(G, G + P * P), c(B, B + 2 * P);
    for (int p = 0; p < P; ++p) gd[p * P + p] += kRidge * tr / P;
    if (!chol_solve_small(P, gd.data(), c.data(), 2)) return NAN;
    double btc = 0.0, ctc = 0.0;
    for (int e = 0; e < 2 * P; ++e) { btc += B[e] * c[e]; ctc += c[e] * c[e]; }
    MpnnDxCoef k{};
    for (int p = 0; p < P; ++p)
        for (int q = 0; q < P; ++q)
            k.p[p * P + q] = 2.0 * (c[p * 2] * c[q * 2] + c[p * 2 + 1] * c[q * 2 + 1]) +
                             (p == q ? 2.0 * kRidge / P * ctc : 0.0);
    for (int t = 0; t < 2; ++t)
        for (int q = 0; q < P; ++q) k.q[t * P + q] = -2.0 * c[q * 2 + t];
    mpnn_dx(n, M.x, vt, k, M.dx, s);
    return vv - btc;
}

// Backward pass from M.dx; gradient (FP64, packed like the parameters) in M.grad.
void mpnn_backward(Mpnn& M, const DevLaplacian& L, const float* f, hipStream_t s) {
    const int64_t n = L.n;
    double* ws = ws_get<double>(kWsGram, gram_workspace_doubles(n, kHid, kHid));
    double* gr = M.grad.get();
    gram_f32(n, M.h[2], kHid, kHid, M.dx, kNout, kNout, gr + M.lay.off[9], ws, s);
    block_mul_f32(n, M.dx, kNout, kNout, M.wot, kHid, nullptr, 0, 0, nullptr, M.u, kHid, s);
    float* dz = M.dz0.get();
    float* dzn = M.dz1.get();
    mpnn_dtanh(n * kHid, M.u, M.h[2], dz, s);
    for (int l = 2; l >= 0; --l) {
        const int din = l == 0 ? kNf : kHid;
        const float* hin = l == 0 ? f : M.h[l - 1].get();
        gram_f32(n, hin, din, din, dz, kHid, kHid, gr + M.lay.off[3 * l], ws, s);
        gram_f32(n, M.av[l], din, din, dz, kHid, kHid, gr + M.lay.off[3 * l + 1], ws, s);
        gram_f32(n, M.ones, 1, 1, dz, kHid, kHid, gr + M.lay.off[3 * l + 2], ws, s);
        if (l == 0) break;
        // dZ_{l-1} = (dZ Ws^T + A_hat (dZ Wn^T)) * (1 - H_{l-1}^2)
        mpnn_gemm(n, kHid, kHid, dz, M.w[l][3], nullptr, nullptr, nullptr, nullptr, 0, nullptr, M.u, s);
        ahat(L, M.u, M.t, kHid, s);
        mpnn_gemm(n, kHid, kHid, dz, M.w[l][2], nullptr, nullptr, nullptr, M.t, 2, M.h[l - 1], dzn, s);
        std::swap(dz, dzn);
    }
}

double mpnn_train_step(Mpnn& M, const DevLaplacian& L, const float* f, const double* vt, float lr,
                       hipStream_t s) {
    mpnn_forward(M, L, f, s);
    const double loss = mpnn_loss(M, L.n, vt, s);
    mpnn_backward(M, L, f, s);
    mpnn_adam(M.lay.total, M.p, M.grad, M.m, M.v, lr, 0.9f, 0.999f, 1e-8f, ++M.step, s);
    M.packed = false;
    return loss;
}

}  // namespace si

struct si_mpnn {
    si::Mpnn m;
    int device = 0;
};

namespace {

template <typename F>
si_status wguard(si_context* ctx, F&& f) {
    try {
        if (ctx) SI_HIP(hipSetDevice(ctx->c.device));
        f();
        if (ctx) ctx->c.last_error.clear();
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

extern "C" {

int64_t si_mpnn_param_count(void) { return si::layout().total; }

si_status si_mpnn_create(si_context* ctx, uint64_t seed, si_mpnn** out) {
    if (!ctx || !out) return SI_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    si_mpnn* m = nullptr;
    const si_status st = wguard(ctx, [&] {
        m = new si_mpnn();
        m->device = ctx->c.device;
        si::mpnn_init(m->m, seed, ctx->c.stream);
    });
    if (st != SI_OK) { delete m; return st; }
    *out = m;
    return SI_OK;
}

void si_mpnn_destroy(si_mpnn* m) {
    if (!m) return;
    hipSetDevice(m->device);
    delete m;
}

si_status si_mpnn_get_params(si_context* ctx, const si_mpnn* m, float* out, int64_t count) {
    if (!ctx || !m || !out) return SI_ERR_INVALID_ARGUMENT;
    return wguard(ctx, [&] {
        SI_CHECK(count == m->m.lay.total, si::kErrInvalidArgument, "mpnn: parameter count mismatch");
        si::to_host(out, m->m.p.get(), count, ctx->c.stream);
    });
}

si_status si_mpnn_set_params(si_context* ctx, si_mpnn* m, const float* in, int64_t count) {
    if (!ctx || !m || !in) return SI_ERR_INVALID_ARGUMENT;
    return wguard(ctx, [&] {
        SI_CHECK(count == m->m.lay.total, si::kErrInvalidArgument, "mpnn: parameter count mismatch");
        si::to_device(m->m.p.get(), in, count, ctx->c.stream);
        m->m.packed = false;
    });
}

si_status si_mpnn_features(si_context* ctx, const si_laplacian* lap, uint64_t seed, float* f) {
    if (!ctx || !lap || !f) return SI_ERR_INVALID_ARGUMENT;
    return wguard(ctx, [&] {
        const si::DevLaplacian& L = lap->L;
        hipStream_t s = ctx->c.stream;
        si::DevBuf<float> fi((size_t)L.n * 16);
        si::mpnn_features(L, seed, fi, s);
        si::rows_to_external(L, fi.get(), 16, f, 16, 16, s);
        SI_HIP(hipStreamSynchronize(s));
    });
}

si_status si_mpnn_forward(si_context* ctx, si_mpnn* m, const si_laplacian* lap, const float* f,
                          float* x) {
    if (!ctx || !m || !lap || !f || !x) return SI_ERR_INVALID_ARGUMENT;
    return wguard(ctx, [&] {
        const si::DevLaplacian& L = lap->L;
        hipStream_t s = ctx->c.stream;
        si::DevBuf<float> fi((size_t)L.n * 16);
        si::rows_to_internal(L, f, 16, fi.get(), 16, 16, s);
        si::mpnn_forward(m->m, L, fi, s);
        si::rows_to_external(L, m->m.x.get(), si::kMpnnOut, x, si::kMpnnOut, si::kMpnnOut, s);
        SI_HIP(hipStreamSynchronize(s));
    });
}

si_status si_mpnn_gradient(si_context* ctx, si_mpnn* m, const si_laplacian* lap, const float* f,
                           const double* v, double* loss, double* grad) {
    if (!ctx || !m || !lap || !f || !v) return SI_ERR_INVALID_ARGUMENT;
    return wguard(ctx, [&] {
        const si::DevLaplacian& L = lap->L;
        hipStream_t s = ctx->c.stream;
        si::DevBuf<float> fi((size_t)L.n * 16);
        si::DevBuf<double> vi((size_t)L.n * 2);
        si::rows_to_internal(L, f, 16, fi.get(), 16, 16, s);
        si::rows_to_internal(L, v, 2, vi.get(), 2, 2, s);
        si::mpnn_forward(m->m, L, fi, s);
        const double l = si::mpnn_loss(m->m, L.n, vi, s);
        si::mpnn_backward(m->m, L, fi, s);
        if (loss) *loss = l;
        if (grad) si::to_host(grad, m->m.grad.get(), m->m.lay.total, s);
    });
}

si_status si_mpnn_train_step(si_context* ctx, si_mpnn* m, const si_laplacian* lap, const float* f,
                             const double* v, float lr, double* loss) {
    if (!ctx || !m || !lap || !f || !v) return SI_ERR_INVALID_ARGUMENT;
    return wguard(ctx, [&] {
        const si::DevLaplacian& L = lap->L;
        hipStream_t s = ctx->c.stream;
        float* fi = si::ws_get<float>(si::kWsAbiIn, (size_t)L.n * 16);
        double* vi = si::ws_get<double>(si::kWsAbiOut, (size_t)L.n * 2);
        si::rows_to_internal(L, f, 16, fi, 16, 16, s);
        si::rows_to_internal(L, v, 2, vi, 2, 2, s);
        const double l = si::mpnn_train_step(m->m, L, fi, vi, lr, s);
        if (loss) *loss = l;
    });
}

si_status si_mpnn_warm_start(si_context* ctx, si_mpnn* m, const si_laplacian* lap, uint64_t seed,
                             int32_t n_components, double* block, int32_t* used_net,
                             double* res_net, double* res_random) {
    if (!ctx || !m || !lap || !block) return SI_ERR_INVALID_ARGUMENT;
    return wguard(ctx, [&] {
        SI_CHECK(n_components == 2, si::kErrUnsupported,
                 "warm start: the network predicts n_components = 2");
        const si::DevLaplacian& L = lap->L;
        const int64_t n = L.n;
        const int k = 3;
        const int W = 1 + si::kMpnnOut * si::kProbeSets;
        const int WP = 32;  // the candidate block stored 32 wide: one SpMM launch
        hipStream_t s = ctx->c.stream;
        // candidate subspaces: [sqrt_deg | X(probe set 0) | X(set 1) | X(set 2)] and
        // the solver's own random start [sqrt_deg | N(0,1)]
        si::DevBuf<float> fi((size_t)n * 16 * si::kProbeSets);
        si::DevBuf<double> bn((size_t)n * WP), br((size_t)n * k);
        SI_HIP(hipMemsetAsync(bn.get(), 0, sizeof(double) * (size_t)n * WP, s));
        si::set_column_f64(bn, WP, n, 0, L.sqrt_deg, 1.0, s);
        si::mpnn_features_sets(L, seed, si::kProbeSets, fi, s);
        for (int set = 0; set < si::kProbeSets; ++set) {
            si::mpnn_forward(m->m, L, fi.get() + (size_t)set * n * 16, s);
            si::convert_f32_to_f64(m->m.x, m->m.xd, n * si::kMpnnOut, s);
            si::copy_cols_f64(m->m.xd, si::kMpnnOut, bn.get() + 1 + set * si::kMpnnOut, WP, n,
                              si::kMpnnOut, s);
        }
        si::fill_normal_f64(br, n, k, k, seed, si::kStreamLobpcgInit, s, L.perm);
        si::set_column_f64(br, k, n, 0, L.sqrt_deg, 1.0, s);
        double ev[3], rs[3];
        // Rayleigh-Ritz of a subspace basis that need not be orthonormal:
        // (B^T L B) c = theta (B^T B) c; the k lowest Ritz vectors go to out,
        // judged by their FP64 residuals
        si::DevBuf<double> lb((size_t)n * WP), gg((size_t)W * 2 * W), th(W), cm((size_t)W * W),
            csub((size_t)W * k), gws(si::gram_workspace_doubles(n, W, 2 * W)), on((size_t)n * k),
            orr((size_t)n * k);
        si::DevBuf<int> stt(1);
        // b: w basis columns stored ld wide (columns past w are zero)
        auto judge = [&](const double* b, int w, int ld, double* out) -> double {
            si::spmm_f64(L, b, lb, ld, ld, 1.0, 0.0, 0.0, nullptr, s);
            si::gram2_f64(n, b, ld, w, b, ld, w, lb, ld, w, gg, gws, s);
            si::small_gen_eig_scaled(w, gg.get() + w, 2 * w, gg, 2 * w, th, cm, stt, s);
            int st = 0;
            si::to_host(&st, stt.get(), 1, s);
            if (st != 0) return INFINITY;
            si::small_copy_sub(w, k, cm, w, csub, k, 1.0, s);
            si::block_mul_f64(n, b, ld, w, csub, k, nullptr, 0, 0, nullptr, out, k, s);
            si::to_host(ev, th.get(), k, s);
            si::fp64_residuals(L, out, k, ev, rs, s);
            double r = 0.0;
            for (double x : rs) r = std::isnan(x) ? INFINITY : std::max(r, x);
            return r;
        };
        const double rn = judge(bn, W, WP, on), rr = judge(br, k, k, orr);
        const bool net = rn < rr;
        si::rows_to_external(L, net ? on.get() : orr.get(), k, block, k, k, s);
        SI_HIP(hipStreamSynchronize(s));
        if (used_net) *used_net = net ? 1 : 0;
        if (res_net) *res_net = rn;
        if (res_random) *res_random = rr;
    });
}

}  // extern "C"
