// k_order.hip -- the clustering locality order of a connected graph.
//
// The 16-column FP32 SpMM is bound by the L2 misses of its x-row gathers, so
// its time is set by how often a row's neighbours are stored near it.  On
// clustered k-NN graphs the best order found groups the rows by cluster:
//
//   1. a short random walk: X <- (I - L) X for 10 steps on 16 columns
//      (N(0,1) start from the Philox stream kStreamOrder), the trivial
//      direction D^{1/2} 1 projected out and the block re-orthonormalised
//      (Cholesky-QR) every 5 steps -- a subspace iteration toward the top
//      of the spectrum of I - L, where nodes of one cluster have similar rows;
//   2. the rows normalised to unit length and clustered by k-means (128
//      centroids, Lloyd iterations; each point's nearest centroid in FP32,
//      ties to the lower index; centroid sums in 64-bit fixed point, so they
//      do not depend on the order of the atomic additions);
//   3. a stable radix sort of the node ids by cluster: each cluster's rows
//      stay in node order.
//
// Every step is deterministic, so one graph always gets the same order.  The
// caller gives the operator in the input node order with its values built.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "si_host.hpp"

namespace si {

namespace {

constexpr int kE = 16;         // embedding columns
constexpr int kSteps = 10;     // walk steps
constexpr int kQrEvery = 5;    // re-orthonormalise every kQrEvery steps
constexpr int kK = 128;        // clusters
constexpr int kIters = 8;      // Lloyd iterations
constexpr float kFix = 1073741824.0f;  // 2^30: fixed-point scale of |y| <= 1
constexpr uint64_t kOrderSeed = 0x5eedc1u;

// y[v] = x[v] / max(|x[v]|, 1e-30) (rows of 16 floats)
__global__ void k_ord_normalize(int64_t n, const float4* __restrict__ x, float4* __restrict__ y) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    float4 a[4];
    float s = 0.f;
    for (int q = 0; q < 4; ++q) {
        a[q] = x[v * 4 + q];
        s += a[q].x * a[q].x + a[q].y * a[q].y + a[q].z * a[q].z + a[q].w * a[q].w;
    }
    const float r = 1.0f / fmaxf(sqrtf(s), 1e-30f);
    for (int q = 0; q < 4; ++q) y[v * 4 + q] = make_float4(a[q].x * r, a[q].y * r, a[q].z * r, a[q].w * r);
}

// initial centroids: the rows at (k + 1/2) n / K
__global__ void k_ord_seed(int64_t n, const float* __restrict__ y, float* __restrict__ cent) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kK * kE) return;
    const int k = t / kE, j = t % kE;
    const int64_t row = ((2 * (int64_t)k + 1) * n) / (2 * kK);
    cent[t] = y[row * kE + j];
}

// nearest centroid of every point; with `accumulate`, the fixed-point sums and
// counts of the points of every cluster (per-workgroup sums in LDS, one global
// addition per workgroup and entry)
__global__ __launch_bounds__(256) void k_ord_assign(int64_t n, const float* __restrict__ y,
                                                    const float* __restrict__ cent,
                                                    int32_t* __restrict__ lab, int accumulate,
                                                    unsigned long long* __restrict__ sums,
                                                    unsigned long long* __restrict__ counts) {
    __shared__ float c[kK * kE];
    __shared__ unsigned long long ls[kK * kE];
    __shared__ unsigned long long lc[kK];
    for (int t = threadIdx.x; t < kK * kE; t += blockDim.x) {
        c[t] = cent[t];
        ls[t] = 0ull;
    }
    for (int t = threadIdx.x; t < kK; t += blockDim.x) lc[t] = 0ull;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        float p[kE];
        for (int j = 0; j < kE; ++j) p[j] = y[v * kE + j];
        float best = INFINITY;
        int bk = 0;
        for (int k = 0; k < kK; ++k) {
            float d = 0.f;
            for (int j = 0; j < kE; ++j) {
                const float t = p[j] - c[k * kE + j];
                d += t * t;
            }
            if (d < best) {
                best = d;
                bk = k;
            }
        }
        lab[v] = bk;
        if (accumulate) {
            for (int j = 0; j < kE; ++j)
                atomicAdd(&ls[bk * kE + j], (unsigned long long)(long long)(p[j] * kFix));
            atomicAdd(&lc[bk], 1ull);
        }
    }
    if (!accumulate) return;
    __syncthreads();
    for (int t = threadIdx.x; t < kK * kE; t += blockDim.x)
        if (ls[t]) atomicAdd(&sums[t], ls[t]);
    for (int t = threadIdx.x; t < kK; t += blockDim.x)
        if (lc[t]) atomicAdd(&counts[t], lc[t]);
}

// Lloyd update: each non-empty cluster's centroid = its fixed-point sum over
// its count (FP64 division, rounded to FP32); an empty cluster keeps its
// centroid.  One thread per centroid entry.
__global__ void k_ord_update(const unsigned long long* __restrict__ sums,
                             const unsigned long long* __restrict__ counts,
                             float* __restrict__ cent) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kK * kE) return;
    const unsigned long long cnt = counts[t / kE];
    if (!cnt) return;
    const double sum = (double)(long long)sums[t];
    cent[t] = (float)(sum / ((double)kFix * (double)cnt));
}

__global__ void k_ord_iota(int64_t n, int32_t* __restrict__ a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (int32_t)i;
}

__global__ void k_ord_inverse(int64_t n, const int32_t* __restrict__ perm, int32_t* __restrict__ iperm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) iperm[perm[i]] = (int32_t)i;
}

inline dim3 grid_of(int64_t threads) { return dim3((unsigned)ceil_div(threads, 256)); }

// Upper T with (X T)^T (X T) = I from G = X^T X (16 x 16), by Cholesky on the
// host; false when G is not positive definite.
bool chol_inv_upper(const std::vector<double>& g, std::vector<double>& t) {
    const int m = kE;
    std::vector<double> l((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = g[(size_t)i * m + j];
            for (int q = 0; q < j; ++q) s -= l[(size_t)i * m + q] * l[(size_t)j * m + q];
            if (i == j) {
                if (!(s > 0.0)) return false;
                l[(size_t)i * m + i] = std::sqrt(s);
            } else {
                l[(size_t)i * m + j] = s / l[(size_t)j * m + j];
            }
        }
    // T = L^-T: solve L^T T = I column by column (T upper triangular)
    t.assign((size_t)m * m, 0.0);
    for (int c = 0; c < m; ++c)
        for (int i = c; i >= 0; --i) {
            double s = i == c ? 1.0 : 0.0;
            for (int q = i + 1; q <= c; ++q) s -= l[(size_t)q * m + i] * t[(size_t)q * m + c];
            t[(size_t)i * m + c] = s / l[(size_t)i * m + i];
        }
    return true;
}

}  // namespace

void cluster_order_device(const DevLaplacian& L, int32_t* perm, int32_t* iperm, hipStream_t s) {
    const int64_t n = L.n;
    SI_CHECK(n >= 2 * kK && n < (int64_t)INT32_MAX, kErrInternal, "cluster order: size out of range");
    const size_t nm = (size_t)n * kE;
    StreamBuf<float> xa(nm, s), xb(nm, s), v0(n, s);
    StreamBuf<double> small(kE * kE * 2 + 64, s);
    StreamBuf<double> ws(gram_workspace_doubles(n, kE, kE), s);
    double* g = small.get();            // 16 x 16 Gram / transforms
    double* c = g + kE * kE;            // second small matrix
    std::vector<double> hg(kE * kE), ht, hid(kE * kE, 0.0);
    for (int j = 0; j < kE; ++j) hid[(size_t)j * kE + j] = 1.0;
    convert_f64_to_f32(L.sqrt_deg, v0.get(), n, s);
    double vv = 0.0;
    gram_f32(n, v0.get(), 1, 1, v0.get(), 1, 1, g, ws.get(), s);
    to_host(&vv, g, 1, s);
    float* x = xa.get();
    float* y = xb.get();
    // project out D^{1/2} 1, then X <- X T with T = chol(X^T X)^-T
    auto orthonormalise = [&] {
        if (vv > 0.0) {
            std::vector<double> h(kE);
            gram_f32(n, v0.get(), 1, 1, x, kE, kE, g, ws.get(), s);
            to_host(h.data(), g, kE, s);
            std::vector<double> up((size_t)kE * kE + kE);
            for (int j = 0; j < kE; ++j) up[j] = -h[j] / vv;
            std::copy(hid.begin(), hid.end(), up.begin() + kE);
            to_device(c, up.data(), up.size(), s);
            block_mul_f32(n, v0.get(), 1, 1, c, kE, x, kE, kE, c + kE, y, kE, s);
            std::swap(x, y);
        }
        gram_f32(n, x, kE, kE, x, kE, kE, g, ws.get(), s);
        to_host(hg.data(), g, hg.size(), s);
        if (!chol_inv_upper(hg, ht)) return;  // keep the block as it is
        to_device(c, ht.data(), ht.size(), s);
        block_mul_f32(n, x, kE, kE, c, kE, nullptr, 0, 0, nullptr, y, kE, s);
        std::swap(x, y);
    };
    fill_normal_rounded_f32(x, n, kE, kE, kOrderSeed, kStreamOrder, s);
    orthonormalise();
    for (int step = 1; step <= kSteps; ++step) {
        spmm_f32(L, x, y, kE, kE, -1.0f, 1.0f, 0.0f, nullptr, s);  // y = (I - L) x
        std::swap(x, y);
        if (step % kQrEvery == 0) orthonormalise();
    }
    hipLaunchKernelGGL(k_ord_normalize, grid_of(n), dim3(256), 0, s, n,
                       reinterpret_cast<const float4*>(x), reinterpret_cast<float4*>(y));
    SI_LAUNCH_CHECK();
    const float* pts = y;
    // k-means
    StreamBuf<float> cent(kK * kE, s);
    StreamBuf<unsigned long long> acc(kK * kE + kK, s);
    StreamBuf<int32_t> lab(n, s), ids(n, s), lab2(n, s);
    hipLaunchKernelGGL(k_ord_seed, grid_of(kK * kE), dim3(256), 0, s, n, pts, cent.get());
    SI_LAUNCH_CHECK();
    const unsigned blocks = (unsigned)std::min<int64_t>(1024, ceil_div(n, 256));
    // Lloyd iterations on the device, no host round trip in between
    for (int it = 0; it < kIters; ++it) {
        SI_HIP(hipMemsetAsync(acc.get(), 0, sizeof(unsigned long long) * acc.n, s));
        hipLaunchKernelGGL(k_ord_assign, dim3(blocks), dim3(256), 0, s, n, pts, cent.get(),
                           lab.get(), 1, acc.get(), acc.get() + kK * kE);
        hipLaunchKernelGGL(k_ord_update, grid_of(kK * kE), dim3(256), 0, s, acc.get(),
                           acc.get() + kK * kE, cent.get());
        SI_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_ord_assign, dim3(blocks), dim3(256), 0, s, n, pts, cent.get(), lab.get(),
                       0, acc.get(), acc.get() + kK * kE);
    hipLaunchKernelGGL(k_ord_iota, grid_of(n), dim3(256), 0, s, n, ids.get());
    SI_LAUNCH_CHECK();
    // node ids ordered by cluster, node order inside each (stable radix sort)
    size_t tb = 0;
    SI_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, lab.get(), lab2.get(), ids.get(), perm,
                                              (int)n, 0, 8, s));
    StreamBuf<char> tmp(tb, s);
    SI_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.get(), tb, lab.get(), lab2.get(), ids.get(), perm,
                                              (int)n, 0, 8, s));
    hipLaunchKernelGGL(k_ord_inverse, grid_of(n), dim3(256), 0, s, n, perm, iperm);
    SI_LAUNCH_CHECK();
    SI_HIP(hipStreamSynchronize(s));
}

}  // namespace si
