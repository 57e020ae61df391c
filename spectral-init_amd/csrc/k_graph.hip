// k_graph.hip -- the fuzzy k-NN graph W (the input of spectral_init) on the
// device: the steps of the reference's fixture pipeline
// (tests/generate_fixtures.py:70-182, umap-learn 0.5.11):
//
//   k_knn            exact k-NN, self excluded (generate_fixtures.py:70-93).
//                    One query per lane, Q queries per thread; candidate points
//                    staged through LDS in tiles of 64 (in the input precision);
//                    FP64 sums of squared differences in dimension order
//                    (fused), top-k kept sorted in registers, ties to the
//                    smaller index.  Bound by the LDS broadcast reads and the
//                    FP64 VALU work (2-3 instructions per dimension and pair).
//   k_smooth_knn     umap smooth_knn_dist, one row per thread (float32 bisection).
//   k_membership     umap compute_membership_strengths, one entry per thread.
//   fuzzy union      W = A + A^T - A o A^T: every directed entry is written
//                    twice (as (i,j) of A and as (j,i) of A^T), the 2nk keys
//                    row*n+col are radix-sorted (hipCUB), equal keys merged
//                    (at most two per key), zeros dropped, then pruned below
//                    max(W)/n_epochs and compacted into CSR.
//
// exp: the float32 exponentials use the method of glibc's expf (ARM
// optimized-routines: 2^(k/32) from a 32-entry table, a cubic in the
// remainder, all in double, one rounding to float), so membership values and
// sigma bisections round as the reference fixtures (generated on glibc) do.
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "si_graph.hpp"

namespace si {

namespace {

// 2^(i/32) as double bit patterns minus (i << 47); computed on the host once.
__constant__ uint64_t c_exp2_tab[32];

__device__ inline float expf_tab(float x) {
#pragma clang fp contract(off)
    // x <= 0 on every call site; below -103.97 the result underflows to 0
    if (x < -0x1.9fe368p6f) return 0.0f;
    const double N = 32.0;
    const double invln2N = 0x1.71547652b82fep+0 * N;
    const double shift = 0x1.8p+52;
    const double c0 = 0x1.c6af84b912394p-5 / N / N / N;
    const double c1 = 0x1.ebfce50fac4f3p-3 / N / N;
    const double c2 = 0x1.62e42ff0c52d6p-1 / N;
    const double z = invln2N * (double)x;
    double kd = z + shift;
    const uint64_t ki = (uint64_t)__double_as_longlong(kd);
    kd -= shift;
    const double r = z - kd;
    const uint64_t t = c_exp2_tab[ki % 32] + (ki << 47);
    const double s = __longlong_as_double((long long)t);
    const double zz = fma(c0, r, c1);
    const double r2 = r * r;
    double y = fma(c2, r, 1.0);
    y = fma(zz, r2, y);
    return (float)(y * s);
}

template <typename T>
__global__ void k_pad_points(const T* __restrict__ src, int64_t n, int d, int dp,
                             T* __restrict__ dst) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * dp) return;
    const int64_t i = e / dp;
    const int t = (int)(e % dp);
    dst[e] = t < d ? src[i * d + t] : T(0);
}

// Top-KP insertion of (s, j) into the ascending list (bd, bi); strict
// comparisons keep the earlier index first on equal distances.
template <int KP>
__device__ inline void topk_insert(double (&bd)[KP], int32_t (&bi)[KP], double s, int32_t j) {
#pragma unroll
    for (int p = KP - 1; p > 0; --p) {
        const bool mv = s < bd[p - 1];
        const bool here = !mv && s < bd[p];
        bd[p] = mv ? bd[p - 1] : (here ? s : bd[p]);
        bi[p] = mv ? bi[p - 1] : (here ? j : bi[p]);
    }
    if (s < bd[0]) { bd[0] = s; bi[0] = j; }
}

constexpr int kKnnTile = 64;   // candidates per LDS tile
constexpr int kKnnThreads = 256;

template <typename CT> struct Lds16;
template <> struct Lds16<float> { using type = float4; };
template <> struct Lds16<double> { using type = double2; };

// CT is the storage type of the points (float32 inputs stay float32 in HBM
// and LDS, which halves the LDS bytes of the broadcast reads that bound this
// kernel; every coordinate is widened to FP64 exactly before it is used).
template <int DP, int KP, int Q, typename CT>
__global__ __launch_bounds__(kKnnThreads) void k_knn(const CT* __restrict__ x, int64_t n,
                                                     int k, int32_t* __restrict__ oi,
                                                     float* __restrict__ od) {
    using V = typename Lds16<CT>::type;
    constexpr int VN = 16 / (int)sizeof(CT);  // coordinates per 16-byte LDS read
    static_assert(DP % VN == 0, "padded dimension must fill 16-byte reads");
    __shared__ CT tile[kKnnTile * DP];
    int64_t qi[Q];
    double q[Q][DP];
    double bd[Q][KP];
    int32_t bi[Q][KP];
#pragma unroll
    for (int u = 0; u < Q; ++u) {
        qi[u] = ((int64_t)blockIdx.x * Q + u) * kKnnThreads + threadIdx.x;
        const bool ok = qi[u] < n;
#pragma unroll
        for (int t = 0; t < DP; ++t) q[u][t] = ok ? (double)x[qi[u] * DP + t] : 0.0;
#pragma unroll
        for (int p = 0; p < KP; ++p) { bd[u][p] = INFINITY; bi[u][p] = -1; }
    }
    for (int64_t c0 = 0; c0 < n; c0 += kKnnTile) {
        const int cnt = (int)((n - c0) < kKnnTile ? (n - c0) : kKnnTile);
        __syncthreads();
        for (int e = threadIdx.x; e < kKnnTile * DP; e += kKnnThreads)
            tile[e] = e < cnt * DP ? x[c0 * DP + e] : CT(0);
        __syncthreads();
        for (int c = 0; c < cnt; ++c) {
            double s[Q];
#pragma unroll
            for (int u = 0; u < Q; ++u) s[u] = 0.0;
#pragma unroll
            for (int t = 0; t < DP; t += VN) {
                const V v = *reinterpret_cast<const V*>(&tile[c * DP + t]);
                const CT* vc = reinterpret_cast<const CT*>(&v);
#pragma unroll
                for (int e = 0; e < VN; ++e) {
                    const double cd = (double)vc[e];
#pragma unroll
                    for (int u = 0; u < Q; ++u) {
                        const double df = q[u][t + e] - cd;
                        s[u] = fma(df, df, s[u]);
                    }
                }
            }
            const int64_t j = c0 + c;
#pragma unroll
            for (int u = 0; u < Q; ++u)
                if (s[u] < bd[u][KP - 1] && j != qi[u]) topk_insert<KP>(bd[u], bi[u], s[u], (int32_t)j);
        }
    }
#pragma unroll
    for (int u = 0; u < Q; ++u) {
        if (qi[u] >= n) continue;
#pragma unroll
        for (int p = 0; p < KP; ++p) {
            if (p < k) {
                oi[qi[u] * k + p] = bi[u][p];
                od[qi[u] * k + p] = bi[u][p] >= 0 ? (float)sqrt(bd[u][p]) : INFINITY;
            }
        }
    }
}

// umap smooth_knn_dist (umap 0.5.11; generate_fixtures.py:96-111): float32
// bisection as umap declares it.  umap is compiled with fastmath (no-infs),
// which folds its "hi == inf" test to false, so the doubling step becomes
// (lo + inf) / 2 = inf: a row whose first sum is at or below log2(k) ends with
// sigma = inf, exactly as the fixtures hold it; fastmath's reciprocal form
// d * (1 / mid) is kept too.
__global__ void k_smooth_knn(const float* __restrict__ dist, int64_t n, int k, float target,
                             float mean_all, float* __restrict__ sig, float* __restrict__ rho_out) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* di = dist + i * k;
    float rho = 0.0f;
    for (int j = 0; j < k; ++j)
        if (di[j] > 0.0f) { rho = di[j]; break; }
    float lo = 0.0f, hi = INFINITY, mid = 1.0f;
    for (int it = 0; it < 64; ++it) {
        float psum = 0.0f;
        const float rm = 1.0f / mid;
        for (int j = 1; j < k; ++j) {
            const float dd = di[j] - rho;
            psum += dd > 0.0f ? expf_tab(-dd * rm) : 1.0f;
        }
        if (fabsf(psum - target) < 1e-5f) break;
        if (psum > target) hi = mid;
        else lo = mid;
        mid = (lo + hi) / 2.0f;
    }
    float sigma = mid;
    if (rho > 0.0f) {
        float m = 0.0f;
        for (int j = 0; j < k; ++j) m += di[j];
        m /= (float)k;
        if ((double)sigma < 1e-3 * (double)m) sigma = (float)(1e-3 * (double)m);
    } else if ((double)sigma < 1e-3 * (double)mean_all) {
        sigma = (float)(1e-3 * (double)mean_all);
    }
    sig[i] = sigma;
    rho_out[i] = rho;
}

// umap compute_membership_strengths (generate_fixtures.py:127-132), float32.
__global__ void k_membership(const int32_t* __restrict__ idx, const float* __restrict__ dist,
                             const float* __restrict__ sig, const float* __restrict__ rho,
                             int64_t n, int k, float* __restrict__ vals) {
#pragma clang fp contract(off)
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n * k) return;
    const int64_t i = p / k;
    const int32_t j = idx[p];
    float v;
    if (j < 0 || j == i) v = 0.0f;
    else if (dist[p] - rho[i] <= 0.0f || sig[i] == 0.0f) v = 1.0f;
    else v = expf_tab(-((dist[p] - rho[i]) / sig[i]));
    vals[p] = v;
}

// Every directed entry (i, j, v) of A becomes the key i*n+j carrying (v, 0)
// and the key j*n+i carrying (0, v); missing or zero entries get the
// sentinel key n*n, which sorts after every real key.
__global__ void k_union_keys(const int32_t* __restrict__ idx, const float* __restrict__ vals,
                             int64_t n, int k, uint64_t* __restrict__ key,
                             uint64_t* __restrict__ pay) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n * k) return;
    const int64_t i = p / k;
    const int32_t j = idx[p];
    const float v = vals[p];
    const uint64_t sentinel = (uint64_t)n * (uint64_t)n;
    if (j < 0 || v == 0.0f) {
        key[2 * p] = sentinel;
        key[2 * p + 1] = sentinel;
        pay[2 * p] = 0;
        pay[2 * p + 1] = 0;
        return;
    }
    const uint32_t vb = __float_as_uint(v);
    key[2 * p] = (uint64_t)i * (uint64_t)n + (uint64_t)j;
    pay[2 * p] = (uint64_t)vb;                 // A entry in the low word
    key[2 * p + 1] = (uint64_t)j * (uint64_t)n + (uint64_t)i;
    pay[2 * p + 1] = (uint64_t)vb << 32;       // A^T entry in the high word
}

__global__ void k_union_heads(const uint64_t* __restrict__ key, int64_t m, uint64_t sentinel,
                              int32_t* __restrict__ head) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    head[t] = key[t] != sentinel && (t == 0 || key[t] != key[t - 1]) ? 1 : 0;
}

// w = (a + a^T) - a * a^T in float32, in the order scipy evaluates
// A + A.T - A.multiply(A.T) (generate_fixtures.py:155).
__global__ void k_union_combine(const uint64_t* __restrict__ key, const uint64_t* __restrict__ pay,
                                const int32_t* __restrict__ head, const int32_t* __restrict__ pos,
                                int64_t m, uint64_t* __restrict__ ukey, float* __restrict__ uw) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m || !head[t]) return;
    uint64_t pl = pay[t];
    if (t + 1 < m && key[t + 1] == key[t]) pl |= pay[t + 1];
    const float a = __uint_as_float((uint32_t)pl);
    const float at = __uint_as_float((uint32_t)(pl >> 32));
    const float s = a + at;
    const float pr = a * at;
    ukey[pos[t]] = key[t];
    uw[pos[t]] = s - pr;
}

__global__ void k_union_keep(const float* __restrict__ uw, int64_t u, float thr, int prune,
                             int32_t* __restrict__ keep) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= u) return;
    const float w = uw[t];
    keep[t] = (w != 0.0f && !(prune && w < thr)) ? 1 : 0;
}

__global__ void k_union_emit(const uint64_t* __restrict__ ukey, const float* __restrict__ uw,
                             const int32_t* __restrict__ keep, const int32_t* __restrict__ kpos,
                             int64_t u, int64_t n, int32_t* __restrict__ ix, float* __restrict__ w,
                             int32_t* __restrict__ rowcnt) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= u || !keep[t]) return;
    const uint64_t kk = ukey[t];
    const int64_t r = (int64_t)(kk / (uint64_t)n);
    ix[kpos[t]] = (int32_t)(kk % (uint64_t)n);
    w[kpos[t]] = uw[t];
    atomicAdd(&rowcnt[r], 1);
}

__global__ void k_indptr(const int32_t* __restrict__ start, int64_t n, int64_t nnz,
                         int64_t* __restrict__ ip) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ip[i] = start[i];
    if (i == n) ip[n] = nnz;
}

// Sum of m floats in double, fixed order: block b sums the strided elements
// b*256 + t + j*gridDim*256 per thread, then a tree; a second launch with one
// block sums the block partials the same way.
template <typename T>
__global__ void k_sum_to_f64(const T* __restrict__ v, int64_t m, double* __restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < m; t += stride)
        s += (double)v[t];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

void upload_exp_table() {
    static std::mutex mu;
    static bool done[64] = {false};
    int dev = 0;
    SI_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if (dev < 64 && done[dev]) return;
    uint64_t tab[32];
    for (int i = 0; i < 32; ++i) {
        const double v = std::exp2(i / 32.0);
        uint64_t b;
        std::memcpy(&b, &v, 8);
        tab[i] = b - ((uint64_t)i << 47);
    }
    SI_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_exp2_tab), tab, sizeof(tab)));
    if (dev < 64) done[dev] = true;
}

template <int DP, int KP, typename CT>
void launch_knn_q(const CT* x, int64_t n, int k, int32_t* oi, float* od, hipStream_t s) {
    // two queries per thread halve the LDS reads per distance where the
    // registers allow it (DP <= 16, k <= 16; above, occupancy drops to one wave)
    constexpr int Q = (DP <= 16 && KP <= 16) ? 2 : 1;
    const int64_t blocks = ceil_div(n, (int64_t)kKnnThreads * Q);
    hipLaunchKernelGGL((k_knn<DP, KP, Q, CT>), dim3(blocks), dim3(kKnnThreads), 0, s, x, n, k,
                       oi, od);
    SI_LAUNCH_CHECK();
}

template <int DP, typename CT>
void launch_knn_k(const CT* x, int64_t n, int k, int32_t* oi, float* od, hipStream_t s) {
    if (k <= 8) launch_knn_q<DP, 8>(x, n, k, oi, od, s);
    else if (k <= 16) launch_knn_q<DP, 16>(x, n, k, oi, od, s);
    else if (k <= 32) launch_knn_q<DP, 32>(x, n, k, oi, od, s);
    else launch_knn_q<DP, 64>(x, n, k, oi, od, s);
}

int padded_dim(int64_t d, int elem_bytes) {
    const int dps[] = {2, 4, 8, 16, 32, 48, 64};
    const int vn = 16 / elem_bytes;
    for (int p : dps)
        if (d <= p && p % vn == 0) return p;
    return -1;
}

// Pads the points to dp coordinates with zeros (which leave every distance
// unchanged) and runs the k-NN kernel of that width.
template <typename CT>
void knn_typed(const CT* pts, int64_t n, int64_t d, int64_t k, int32_t* idx, float* dist,
               hipStream_t s) {
    const int dp = padded_dim(d, (int)sizeof(CT));
    SI_CHECK(dp > 0, kErrUnsupported, "knn: dimension > 64 is not supported");
    DevBuf<CT> xp((size_t)n * dp);
    const int64_t tot = n * dp;
    hipLaunchKernelGGL((k_pad_points<CT>), dim3(ceil_div(tot, 256)), dim3(256), 0, s, pts, n,
                       (int)d, dp, xp.get());
    SI_LAUNCH_CHECK();
    switch (dp) {
        case 2: if constexpr (sizeof(CT) == 8) launch_knn_k<2>(xp.get(), n, (int)k, idx, dist, s); break;
        case 4: launch_knn_k<4>(xp.get(), n, (int)k, idx, dist, s); break;
        case 8: launch_knn_k<8>(xp.get(), n, (int)k, idx, dist, s); break;
        case 16: launch_knn_k<16>(xp.get(), n, (int)k, idx, dist, s); break;
        case 32: launch_knn_k<32>(xp.get(), n, (int)k, idx, dist, s); break;
        case 48: launch_knn_k<48>(xp.get(), n, (int)k, idx, dist, s); break;
        default: launch_knn_k<64>(xp.get(), n, (int)k, idx, dist, s); break;
    }
    SI_HIP(hipStreamSynchronize(s));  // xp is freed on return
}

}  // namespace

void knn_exact_device(const void* points_dev, int precision, int64_t n, int64_t d, int64_t k,
                      int32_t* idx, float* dist, hipStream_t s) {
    SI_CHECK(n >= 2 && d >= 1 && k >= 1, kErrInvalidArgument, "knn: need n >= 2, d >= 1, k >= 1");
    SI_CHECK(k < n, kErrInvalidArgument, "knn: need k < n (self is excluded)");
    SI_CHECK(k <= 64, kErrUnsupported, "knn: k > 64 is not supported");
    SI_CHECK(n < ((int64_t)1 << 31), kErrUnsupported, "knn: n >= 2^31 is not supported");
    SI_CHECK(d <= 64, kErrUnsupported, "knn: dimension > 64 is not supported");
    if (precision == 0)
        knn_typed<float>(static_cast<const float*>(points_dev), n, d, k, idx, dist, s);
    else
        knn_typed<double>(static_cast<const double*>(points_dev), n, d, k, idx, dist, s);
}

void smooth_knn_device(const float* dist, int64_t n, int64_t k, float* sig, float* rho,
                       hipStream_t s) {
    SI_CHECK(n >= 1 && k >= 2, kErrInvalidArgument, "smooth_knn_dist: need n >= 1, k >= 2");
    upload_exp_table();
    // mean of all distances, for rows without a non-zero distance
    constexpr int kParts = 512;
    DevBuf<double> part(kParts), sum(1);
    hipLaunchKernelGGL((k_sum_to_f64<float>), dim3(kParts), dim3(256), 0, s, dist, n * k,
                       part.get());
    hipLaunchKernelGGL((k_sum_to_f64<double>), dim3(1), dim3(256), 0, s, part.get(),
                       (int64_t)kParts, sum.get());
    SI_LAUNCH_CHECK();
    double h = 0.0;
    SI_HIP(hipMemcpyAsync(&h, sum.get(), sizeof(double), hipMemcpyDeviceToHost, s));
    SI_HIP(hipStreamSynchronize(s));
    const float mean_all = (float)(h / (double)(n * k));
    const float target = (float)std::log2((double)k);
    hipLaunchKernelGGL(k_smooth_knn, dim3(ceil_div(n, 256)), dim3(256), 0, s, dist, n, (int)k,
                       target, mean_all, sig, rho);
    SI_LAUNCH_CHECK();
}

void membership_device(const int32_t* idx, const float* dist, const float* sig, const float* rho,
                       int64_t n, int64_t k, float* vals, hipStream_t s) {
    upload_exp_table();
    hipLaunchKernelGGL(k_membership, dim3(ceil_div(n * k, 256)), dim3(256), 0, s, idx, dist, sig,
                       rho, n, (int)k, vals);
    SI_LAUNCH_CHECK();
}

void fuzzy_union_device(const int32_t* idx, const float* vals, int64_t n, int64_t k,
                        int n_epochs, DevGraph& g, hipStream_t s) {
    SI_CHECK(n >= 1 && k >= 1, kErrInvalidArgument, "fuzzy_union: need n >= 1, k >= 1");
    SI_CHECK(n < ((int64_t)1 << 31), kErrUnsupported, "fuzzy_union: n >= 2^31 is not supported");
    const int64_t m = 2 * n * k;
    SI_CHECK(m < ((int64_t)1 << 31), kErrUnsupported,
             "fuzzy_union: 2 * n * k >= 2^31 entries is not supported");
    const uint64_t sentinel = (uint64_t)n * (uint64_t)n;
    int end_bit = 1;
    while (end_bit < 64 && (sentinel >> end_bit) != 0) ++end_bit;
    DevBuf<uint64_t> key(m), key2(m), pay(m), pay2(m);
    hipLaunchKernelGGL(k_union_keys, dim3(ceil_div(n * k, 256)), dim3(256), 0, s, idx, vals, n,
                       (int)k, key.get(), pay.get());
    SI_LAUNCH_CHECK();
    size_t tb = 0;
    SI_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key.get(), key2.get(), pay.get(),
                                              pay2.get(), (int)m, 0, end_bit, s));
    DevBuf<char> tmp(tb);
    SI_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.get(), tb, key.get(), key2.get(), pay.get(),
                                              pay2.get(), (int)m, 0, end_bit, s));
    // heads of equal-key runs and their compacted positions
    DevBuf<int32_t> head(m), pos(m + 1);
    hipLaunchKernelGGL(k_union_heads, dim3(ceil_div(m, 256)), dim3(256), 0, s, key2.get(), m,
                       sentinel, head.get());
    SI_LAUNCH_CHECK();
    size_t tb2 = 0;
    SI_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb2, head.get(), pos.get() + 1, (int)m, s));
    DevBuf<char> tmp2(tb2);
    SI_HIP(hipMemsetAsync(pos.get(), 0, sizeof(int32_t), s));
    SI_HIP(hipcub::DeviceScan::InclusiveSum(tmp2.get(), tb2, head.get(), pos.get() + 1, (int)m, s));
    int32_t u32 = 0;
    SI_HIP(hipMemcpyAsync(&u32, pos.get() + m, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SI_HIP(hipStreamSynchronize(s));
    const int64_t u = u32;
    if (u == 0) {  // no non-zero membership: a graph without edges
        g.n = n;
        g.nnz = 0;
        g.ip.alloc(n + 1);
        g.ix.alloc(0);
        g.w.alloc(0);
        SI_HIP(hipMemsetAsync(g.ip.get(), 0, sizeof(int64_t) * (n + 1), s));
        SI_HIP(hipStreamSynchronize(s));
        return;
    }
    DevBuf<uint64_t> ukey(u);
    DevBuf<float> uw(u);
    hipLaunchKernelGGL(k_union_combine, dim3(ceil_div(m, 256)), dim3(256), 0, s, key2.get(),
                       pay2.get(), head.get(), pos.get(), m, ukey.get(), uw.get());
    SI_LAUNCH_CHECK();
    // prune threshold max(W) / n_epochs in float32 (generate_fixtures.py:176-178)
    float thr = 0.0f;
    const int prune = n_epochs > 0 && u > 0;
    if (prune) {
        DevBuf<float> mx(1);
        size_t tb3 = 0;
        SI_HIP(hipcub::DeviceReduce::Max(nullptr, tb3, uw.get(), mx.get(), (int)u, s));
        DevBuf<char> tmp3(tb3);
        SI_HIP(hipcub::DeviceReduce::Max(tmp3.get(), tb3, uw.get(), mx.get(), (int)u, s));
        float wmax = 0.0f;
        SI_HIP(hipMemcpyAsync(&wmax, mx.get(), sizeof(float), hipMemcpyDeviceToHost, s));
        SI_HIP(hipStreamSynchronize(s));
        thr = wmax / (float)n_epochs;
    }
    DevBuf<int32_t> keep(u), kpos(u + 1), rowcnt(n), rowstart(n);
    hipLaunchKernelGGL(k_union_keep, dim3(ceil_div(u, 256)), dim3(256), 0, s, uw.get(), u, thr,
                       prune, keep.get());
    SI_LAUNCH_CHECK();
    size_t tb4 = 0;
    SI_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb4, keep.get(), kpos.get() + 1, (int)u, s));
    DevBuf<char> tmp4(tb4);
    SI_HIP(hipMemsetAsync(kpos.get(), 0, sizeof(int32_t), s));
    SI_HIP(hipcub::DeviceScan::InclusiveSum(tmp4.get(), tb4, keep.get(), kpos.get() + 1, (int)u, s));
    int32_t nnz32 = 0;
    SI_HIP(hipMemcpyAsync(&nnz32, kpos.get() + u, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SI_HIP(hipStreamSynchronize(s));
    g.n = n;
    g.nnz = nnz32;
    g.ip.alloc(n + 1);
    g.ix.alloc(g.nnz);
    g.w.alloc(g.nnz);
    SI_HIP(hipMemsetAsync(rowcnt.get(), 0, sizeof(int32_t) * n, s));
    hipLaunchKernelGGL(k_union_emit, dim3(ceil_div(u, 256)), dim3(256), 0, s, ukey.get(), uw.get(),
                       keep.get(), kpos.get(), u, n, g.ix.get(), g.w.get(), rowcnt.get());
    SI_LAUNCH_CHECK();
    size_t tb5 = 0;
    SI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb5, rowcnt.get(), rowstart.get(), (int)n, s));
    DevBuf<char> tmp5(tb5);
    SI_HIP(hipcub::DeviceScan::ExclusiveSum(tmp5.get(), tb5, rowcnt.get(), rowstart.get(), (int)n, s));
    hipLaunchKernelGGL(k_indptr, dim3(ceil_div(n + 1, 256)), dim3(256), 0, s, rowstart.get(), n,
                       g.nnz, g.ip.get());
    SI_LAUNCH_CHECK();
    SI_HIP(hipStreamSynchronize(s));  // the temporaries are freed on return
}

}  // namespace si
