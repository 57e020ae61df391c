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
//                    k > 32 and the queries the filter below cannot certify
//                    take this kernel.
//   k_knn_filter(3), the same k-NN for k <= 32 (float32 or float64 points):
//   k_knn_rerank     matrix-core distances with rigorous error bounds keep a
//                    short candidate list per query, and FP64 distances of
//                    the certified candidates (k_knn's arithmetic) rank them;
//                    the result is k_knn's, bit for bit.
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

#include <algorithm>
#include <atomic>
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
template <int KP, typename D>
__device__ inline void topk_insert(D (&bd)[KP], int32_t (&bi)[KP], D s, int32_t j) {
#pragma unroll
    for (int p = KP - 1; p > 0; --p) {
        const bool mv = s < bd[p - 1];
        const bool here = !mv && s < bd[p];
        bd[p] = mv ? bd[p - 1] : (here ? s : bd[p]);
        bi[p] = mv ? bi[p - 1] : (here ? j : bi[p]);
    }
    if (s < bd[0]) { bd[0] = s; bi[0] = j; }
}

// The same insertion for float lists (callers guarantee s < bd[KP-1]): slot
// p becomes median(bd[p-1], bd[p], s) = max(bd[p-1], min(bd[p], s)), one
// v_med3_f32, and each compare s < bd[p] serves slots p and p + 1.
template <int KP>
__device__ inline void topk_insert_f(float (&bd)[KP], int32_t (&bi)[KP], float s, int32_t j) {
    bool lt[KP];
#pragma unroll
    for (int p = 0; p < KP; ++p) lt[p] = s < bd[p];
#pragma unroll
    for (int p = KP - 1; p > 0; --p) {
        bd[p] = __builtin_amdgcn_fmed3f(bd[p - 1], bd[p], s);
        bi[p] = lt[p - 1] ? bi[p - 1] : (lt[p] ? j : bi[p]);
    }
    if (lt[0]) { bd[0] = s; bi[0] = j; }
}

constexpr int kKnnTile = 64;   // candidates per LDS tile
constexpr int kKnnThreads = 256;

template <typename CT> struct Lds16;
template <> struct Lds16<float> { using type = float4; };
template <> struct Lds16<double> { using type = double2; };

// CT is the storage type of the points (float32 inputs stay float32 in HBM
// and LDS, which halves the LDS bytes of the broadcast reads that bound this
// kernel; every coordinate is widened to FP64 exactly before it is used).
// With qlist set, the queries are qlist[0, nq) instead of 0 .. n-1.  With
// gridDim.y = S > 1, workgroup (x, y) scans the candidates [y n / S, (y + 1) n
// / S) and writes its sorted top KP (squared distance, index) to
// pd/pi[(slot * S + y) * KP ...]; k_knn_merge ranks the S lists.
template <int DP, int KP, int Q, typename CT>
__global__ __launch_bounds__(kKnnThreads) void k_knn(const CT* __restrict__ x, int64_t n,
                                                     int k, int32_t* __restrict__ oi,
                                                     float* __restrict__ od,
                                                     const int32_t* __restrict__ qlist,
                                                     int64_t nq, double* __restrict__ pd,
                                                     int32_t* __restrict__ pi) {
    using V = typename Lds16<CT>::type;
    constexpr int VN = 16 / (int)sizeof(CT);  // coordinates per 16-byte LDS read
    static_assert(DP % VN == 0, "padded dimension must fill 16-byte reads");
    __shared__ CT tile[kKnnTile * DP];
    int64_t qi[Q], sl[Q];
    double q[Q][DP];
    double bd[Q][KP];
    int32_t bi[Q][KP];
    const int64_t cb = n * blockIdx.y / gridDim.y, ce = n * (blockIdx.y + 1) / gridDim.y;
#pragma unroll
    for (int u = 0; u < Q; ++u) {
        const int64_t slot = ((int64_t)blockIdx.x * Q + u) * kKnnThreads + threadIdx.x;
        sl[u] = slot;
        qi[u] = slot < nq ? (qlist ? (int64_t)qlist[slot] : slot) : n;
        const bool ok = qi[u] < n;
#pragma unroll
        for (int t = 0; t < DP; ++t) q[u][t] = ok ? (double)x[qi[u] * DP + t] : 0.0;
#pragma unroll
        for (int p = 0; p < KP; ++p) { bd[u][p] = INFINITY; bi[u][p] = -1; }
    }
    for (int64_t c0 = cb; c0 < ce; c0 += kKnnTile) {
        const int cnt = (int)((ce - c0) < kKnnTile ? (ce - c0) : kKnnTile);
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
        if (gridDim.y > 1) {
            const int64_t o = (sl[u] * gridDim.y + blockIdx.y) * KP;
#pragma unroll
            for (int p = 0; p < KP; ++p) {
                pd[o + p] = bd[u][p];
                pi[o + p] = bi[u][p];
            }
            continue;
        }
#pragma unroll
        for (int p = 0; p < KP; ++p) {
            if (p < k) {
                oi[qi[u] * k + p] = bi[u][p];
                od[qi[u] * k + p] = bi[u][p] >= 0 ? (float)sqrt(bd[u][p]) : INFINITY;
            }
        }
    }
}

// First k of the S ascending (squared distance, index) lists of each query
// slot, ordered by distance, then index: the single-pass result, whose strict
// insertions in index order give that same order.
__global__ void k_knn_merge(const double* __restrict__ pd, const int32_t* __restrict__ pi,
                            int splits, int kp, int64_t n, int k,
                            const int32_t* __restrict__ qlist, int64_t nq,
                            int32_t* __restrict__ oi, float* __restrict__ od) {
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= nq) return;
    const int64_t q = qlist ? (int64_t)qlist[slot] : slot;
    const int64_t base = slot * splits * kp;
    int h[16];
    for (int g = 0; g < splits; ++g) h[g] = 0;
    for (int p = 0; p < k; ++p) {
        int gb = -1;
        double sb = INFINITY;
        int32_t jb = -1;
        for (int g = 0; g < splits; ++g) {
            if (h[g] >= kp) continue;
            const int64_t e = base + (int64_t)g * kp + h[g];
            const int32_t j = pi[e];
            if (j < 0) continue;
            const double sv = pd[e];
            if (gb < 0 || sv < sb || (sv == sb && j < jb)) { gb = g; sb = sv; jb = j; }
        }
        if (gb >= 0) ++h[gb];
        oi[q * k + p] = jb;
        od[q * k + p] = jb >= 0 ? (float)sqrt(sb) : INFINITY;
    }
}

// ------------------------------------------------------------ MFMA filter
// Points (float32 or float64) with k <= kFiltK take two stages that give
// exactly k_knn's result:
//   1. k_knn_filter: a = |q'|^2 + |c'|^2 - 2 q'.c' on the FP32 matrix cores
//      (v_mfma_f32_16x16x4_f32) for the centred, float-rounded points q', c',
//      and the lower bound l = a - E of the exact squared distance d^2, where
//      E = 2 alpha (|q'|^2 + |c'|^2) + eabs bounds |a - d^2| (rounding of the
//      centred coordinates, the norms, the MFMA sums and the epilogue, with
//      a factor 2 to spare; eabs covers flushed subnormal products).  The
//      factors fold into per-point constants, so l = fma(-2, q'.c', nq + nc)
//      costs what a does.  Every lane keeps, for each of its queries, the
//      kFiltL smallest l of the quarter of the candidates that its MFMA
//      output rows cover (the query itself included; stage 2 skips it).
//   2. k_knn_rerank: u = l + 2E >= d^2 for every kept candidate, and U_k, the
//      k-th smallest u of the query, bounds its k-th distance d_k^2.  A true
//      neighbour has l <= d^2 <= U_k, and a list drops a candidate only with
//      l >= its final largest entry; so when every full list ends above U_k
//      all the neighbours are kept, and their exact FP64 distances (k_knn's
//      arithmetic) over the candidates with l <= U_k are ranked by
//      (distance, index).  Every other query goes to k_knn through a list.
typedef float v4f __attribute__((ext_vector_type(4)));

// 16-query tiles per wavefront: 4, or 2 above 32 coordinates (registers)
template <int KS> constexpr int filt_tiles() { return KS > 8 ? 2 : 4; }
constexpr int kFiltK = 32;                     // largest k of the filtered path
constexpr int kFiltL = 16;                     // kept candidates per lane and query
constexpr int kFiltTile = 256;                 // candidates per LDS tile
constexpr int kFiltGroups = 4;                 // lanes sharing one query (one per quarter)
template <int KS> constexpr int filt_queries() { return 4 * 16 * filt_tiles<KS>(); }  // per workgroup
constexpr int kColParts = 2048;

// Column sums of the points (rows of dx stored coordinates, the first d
// used) in FP64, part b over rows [b n / P, (b + 1) n / P); one thread per
// column (d <= 64 = blockDim).
template <typename CT>
__global__ void k_col_partials(const CT* __restrict__ x, int64_t n, int dx, int d,
                               double* __restrict__ part) {
    const int64_t P = gridDim.x;
    const int64_t r0 = n * blockIdx.x / P, r1 = n * (blockIdx.x + 1) / P;
    const int t = threadIdx.x;
    if (t >= d) return;
    double s = 0.0;
    for (int64_t r = r0; r < r1; ++r) s += (double)x[r * dx + t];
    part[(int64_t)blockIdx.x * d + t] = s;
}

// 16 threads per column (1024 threads, dp <= 64) sum strided parts, then
// lane 0 of each column adds the 16 sums in order (a fixed order).
__global__ __launch_bounds__(1024) void k_col_mean(const double* __restrict__ part, int parts,
                                                   int64_t n, int dp,
                                                   double* __restrict__ mean) {
    __shared__ double red[16][64];
    const int t = threadIdx.x & 63, g = threadIdx.x >> 6;
    double s = 0.0;
    if (t < dp)
        for (int b = g; b < parts; b += 16) s += part[(int64_t)b * dp + t];
    red[g][t] = s;
    __syncthreads();
    if (g == 0 && t < dp) {
        double m = 0.0;
        for (int i = 0; i < 16; ++i) m += red[i][t];
        mean[t] = m / (double)n;
    }
}

// xc [n x DP] = float(x - mean) on the first d coordinates (zeros after),
// nrm = float(|xc|^2) (FP64 sum); the largest nrm's bit pattern is max'ed into
// *nmax (non-negative floats order as their unsigned bit patterns, and a NaN
// orders above +inf, so a non-finite input shows).  x has rows of dx.
template <int DP, typename CT>
__global__ void k_center_norms(const CT* __restrict__ x, int64_t n, int dx, int d,
                               const double* __restrict__ mean, float* __restrict__ xc,
                               float* __restrict__ nrm, unsigned* __restrict__ nmax) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float nv = 0.0f;
    if (i < n) {
        double s = 0.0;
#pragma unroll
        for (int t = 0; t < DP; ++t) {
            // float32 rows are stored at the filter width (dx == DP) with zero
            // padding and zero means there, so no per-coordinate test is needed
            // and the row load vectorizes; float64 rows may be narrower (dx = 2)
            float c;
            if constexpr (sizeof(CT) == 4) c = (float)((double)x[i * DP + t] - mean[t]);
            else c = t < d ? (float)((double)x[i * dx + t] - mean[t]) : 0.0f;
            xc[i * DP + t] = c;
            s += (double)c * (double)c;
        }
        nv = (float)s;
        nrm[i] = nv;
    }
    // wavefront max, then the workgroup's (256 threads) in LDS: one atomic per
    // workgroup keeps the contention on *nmax low
    __shared__ unsigned red[4];
    unsigned b = __float_as_uint(nv);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned o = __shfl_xor(b, off);
        b = o > b ? o : b;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned m = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = red[w] > m ? red[w] : m;
        atomicMax(nmax, m);
    }
}

// Stage 1.  A workgroup owns filt_queries() queries of [q0, q1) (16 per
// MFMA tile, filt_tiles() tiles per wavefront) and streams the candidates of its split
// (blockIdx.y of gridDim.y equal ranges of 0 .. n-1) through LDS; the split
// keeps small inputs on enough workgroups to fill the chip.
// Candidates are the A operand (rows), queries the B operand (columns):
// lane l supplies row / column l % 16 and coordinates (l / 16) KS + s of
// k-step s (A and B share that split, so the dot products are whole).  The
// output row each element holds is read off an MFMA of index-valued operands
// at the start; every element's column must be l % 16 (checked).
template <int KS>
__global__ __launch_bounds__(256) void k_knn_filter(const float* __restrict__ xc,
                                                    const float* __restrict__ nrm, int64_t n,
                                                    int64_t q0, int64_t q1,
                                                    float shrink, float eabs, int k,
                                                    float e2a, float e2b,
                                                    int32_t* __restrict__ kidx,
                                                    float* __restrict__ kd2,
                                                    int32_t* __restrict__ layout_err) {
    constexpr int DP = 4 * KS;
    constexpr int kFiltR = filt_tiles<KS>();
    constexpr int DS = DP + 4;  // LDS row stride: a 16-byte shift per row spreads the banks
    __shared__ __align__(16) float tile[kFiltTile * DS];
    __shared__ float tn[kFiltTile];  // (1 - 2 alpha) |c'|^2, +inf past the last candidate
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lq = lane & 15, lg = lane >> 4;

    const v4f zero = {0.0f, 0.0f, 0.0f, 0.0f};
    const v4f pr = __builtin_amdgcn_mfma_f32_16x16x4f32(lg == 0 ? (float)lq : 0.0f,
                                                        lg == 0 ? 1.0f : 0.0f, zero, 0, 0, 0);
    const v4f pc = __builtin_amdgcn_mfma_f32_16x16x4f32(lg == 0 ? 1.0f : 0.0f,
                                                        lg == 0 ? (float)lq : 0.0f, zero, 0, 0, 0);
    int rowof[4];
    bool bad = false;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        rowof[v] = (int)pr[v] & 15;
        bad = bad || (int)pc[v] != lq || (float)rowof[v] != pr[v];
    }
    if (bad) *layout_err = 1;

    const int64_t qb = q0 + ((int64_t)blockIdx.x * 4 + wave) * (16 * kFiltR);
    const int64_t cb = n * blockIdx.y / gridDim.y, ce = n * (blockIdx.y + 1) / gridDim.y;
    float bq[kFiltR][KS], nq[kFiltR];
    int32_t qi[kFiltR];
    float bd[kFiltR][kFiltL];
    int32_t bi[kFiltR][kFiltL];
#pragma unroll
    for (int r = 0; r < kFiltR; ++r) {
        const int64_t q = qb + 16 * r + lq;
        const bool ok = q < q1;
        qi[r] = ok ? (int32_t)q : -1;
#pragma unroll
        for (int s = 0; s < KS; ++s) bq[r][s] = ok ? xc[q * DP + lg * KS + s] : 0.0f;
        nq[r] = ok ? nrm[q] * shrink - eabs : 0.0f;
#pragma unroll
        for (int p = 0; p < kFiltL; ++p) { bd[r][p] = INFINITY; bi[r][p] = -1; }
    }

    float thr[kFiltR];
#pragma unroll
    for (int r = 0; r < kFiltR; ++r) thr[r] = INFINITY;
    const int pk = (k + 1 + 3) / 4 - 1;  // 4 (pk + 1) >= k + 1 entries
    // sub-tile m: dot products of the 16 candidates with the 4 query tiles
    auto dots = [&](int m, v4f (&acc)[kFiltR]) {
        float a[KS];
        const float* src = &tile[(16 * m + lq) * DS + lg * KS];
        if constexpr (KS % 4 == 0) {
#pragma unroll
            for (int s = 0; s < KS; s += 4) {
                const float4 w = *reinterpret_cast<const float4*>(src + s);
                a[s] = w.x; a[s + 1] = w.y; a[s + 2] = w.z; a[s + 3] = w.w;
            }
        } else {
#pragma unroll
            for (int s = 0; s < KS; ++s) a[s] = src[s];
        }
#pragma unroll
        for (int r = 0; r < kFiltR; ++r) acc[r] = zero;
#pragma unroll
        for (int s = 0; s < KS; ++s)
#pragma unroll
            for (int r = 0; r < kFiltR; ++r)
                acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], bq[r][s], acc[r], 0, 0, 0);
    };
    // bounds l of sub-tile m into the kept lists
    auto keep = [&](int m, int64_t c0, const v4f (&acc)[kFiltR]) {
        float cn[4];
        int32_t cj[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = 16 * m + rowof[v];
            cn[v] = tn[row];
            cj[v] = (int32_t)(c0 + row);  // rows past the end have l = +inf
        }
        float lo[kFiltR][4];
        bool any = false;
#pragma unroll
        for (int r = 0; r < kFiltR; ++r)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                lo[r][v] = fmaf(-2.0f, acc[r][v], nq[r] + cn[v]);
                any = any || lo[r][v] < thr[r];
            }
        if (!any) return;  // one branch for the common case
#pragma unroll
        for (int r = 0; r < kFiltR; ++r)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if (lo[r][v] < thr[r]) {  // thr[r] <= bd[r][kFiltL - 1] always
                    topk_insert_f<kFiltL>(bd[r], bi[r], lo[r][v], cj[v]);
                    thr[r] = fminf(thr[r], bd[r][kFiltL - 1]);
                }
    };
    // the threshold shared by the four lanes of a query, as in k_knn_filter3
    // (same argument with this stage's alpha), once per tile
    auto refresh = [&]() {
#pragma unroll
        for (int r = 0; r < kFiltR; ++r) {
            float t = INFINITY;
#pragma unroll
            for (int p = 0; p < kFiltL; ++p) t = p == pk ? bd[r][p] : t;
            t = fmaxf(t, __shfl_xor(t, 16));
            t = fmaxf(t, __shfl_xor(t, 32));
            const float b = t + (qi[r] >= 0 ? fmaf(e2a, nrm[qi[r]], e2b) : 0.0f);
            const float up = b + fabsf(b) * 0x1p-20f + 1e-30f;
            thr[r] = t < INFINITY ? fminf(up, bd[r][kFiltL - 1]) : bd[r][kFiltL - 1];
        }
    };
    for (int64_t c0 = cb; c0 < ce; c0 += kFiltTile) {
        const int cnt = (int)((ce - c0) < kFiltTile ? (ce - c0) : kFiltTile);
        __syncthreads();
        for (int e = threadIdx.x; e < kFiltTile * KS; e += 256) {
            const int row = e / KS, c4 = e - row * KS;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (row < cnt) v = *reinterpret_cast<const float4*>(xc + (c0 + row) * DP + 4 * c4);
            *reinterpret_cast<float4*>(&tile[row * DS + 4 * c4]) = v;
        }
        for (int e = threadIdx.x; e < kFiltTile; e += 256)
            tn[e] = e < cnt ? nrm[c0 + e] * shrink : INFINITY;
        __syncthreads();
        refresh();
        // software pipeline: sub-tile m + 1's MFMAs are issued before sub-tile
        // m's list updates, which then run while the matrix cores work
        const int subs = (cnt + 15) / 16;
        v4f acc[kFiltR];
        dots(0, acc);
        for (int m = 0; m + 1 < subs; ++m) {
            v4f nxt[kFiltR];
            dots(m + 1, nxt);
            keep(m, c0, acc);
#pragma unroll
            for (int r = 0; r < kFiltR; ++r) acc[r] = nxt[r];
        }
        keep(subs - 1, c0, acc);
    }
#pragma unroll
    for (int r = 0; r < kFiltR; ++r) {
        if (qi[r] < 0) continue;
        const int64_t o =
            (((int64_t)(qi[r] - q0) * gridDim.y + blockIdx.y) * kFiltGroups + lg) * kFiltL;
#pragma unroll
        for (int p = 0; p < kFiltL; ++p) {
            kidx[o + p] = bi[r][p];
            kd2[o + p] = bd[r][p];
        }
    }
}

// ------------------------------------------------------------ bf16x3 filter
// The same stage 1 on the bf16 matrix cores (v_mfma_f32_16x16x32_bf16): every
// FP32 value v is split into bf16 parts h = bf16(v), l = bf16(v - h), and a
// product x y is taken as hx hy + hx ly + lx hy (each part product exact in
// the FP32 accumulator; the dropped terms are below 3.02 2^-16 |x y|).  The
// squared norms ride in two spare k-slots: candidate rows are
// [c', nc, 1, 0 ...] and query rows [-2 q', 1, nq, 0 ...] with nc, nq the
// shrunk norms, so the MFMA output is the lower bound l itself and the
// epilogue is one compare per value.  Rows are DB = 32 KB wide (d + 2 <= DB).

// FP32 -> bf16 bits, round to nearest even (finite inputs)
__device__ inline uint16_t bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline void split_bf16(float v, uint16_t& h, uint16_t& l) {
    h = bf16_rne(v);
    const float hv = __uint_as_float((uint32_t)h << 16);
    l = bf16_rne(v - hv);  // v - hv is exact in FP32
}

// Candidate rows (ah, al) and query rows (bh, bl) [n x DB] of the centred
// points xc [n x dp] and their norms: nc = shrink |c'|^2, nq = shrink |q'|^2 - eabs.
__global__ void k_split_rows(const float* __restrict__ xc, const float* __restrict__ nrm,
                             int64_t n, int d, int dp, int db, float shrink, float eabs,
                             uint16_t* __restrict__ ah, uint16_t* __restrict__ al,
                             uint16_t* __restrict__ bh, uint16_t* __restrict__ bl) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * db) return;
    const int64_t i = e / db;
    const int t = (int)(e - i * db);
    float a = 0.0f, b = 0.0f;
    if (t < d) {
        a = xc[i * dp + t];
        b = -2.0f * a;
    } else if (t == d) {
        a = nrm[i] * shrink;
        b = 1.0f;
    } else if (t == d + 1) {
        a = 1.0f;
        b = nrm[i] * shrink - eabs;
    }
    split_bf16(a, ah[e], al[e]);
    split_bf16(b, bh[e], bl[e]);
}

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef uint16_t v8u16 __attribute__((ext_vector_type(8)));

template <int KB>
__global__ __launch_bounds__(256) void k_knn_filter3(
    const uint16_t* __restrict__ ah, const uint16_t* __restrict__ al,
    const uint16_t* __restrict__ bh, const uint16_t* __restrict__ bl, int64_t n, int d, int k,
    const float* __restrict__ nrm, float e2a, float e2b, int64_t q0, int64_t q1,
    int32_t* __restrict__ kidx, float* __restrict__ kd2, int32_t* __restrict__ layout_err) {
    constexpr int DB = 32 * KB;
    constexpr int DS = DB + 8;       // LDS row stride (bf16): spreads the 16-byte reads over the banks
    // candidates per LDS tile, a multiple of the 16-row sub-tile (every
    // sub-tile read stays inside the tile): 256, 128, 80 rows, <= 40 KB of LDS,
    // so LDS does not cap the waves per SIMD
    constexpr int kT = (256 / KB) / 16 * 16;
    static_assert(kT % 16 == 0, "tile rows must be whole sub-tiles");
    // 16-query tiles per wavefront: two keep the registers low enough for three
    // waves per SIMD (four tiles and two waves measured 12 % slower)
    constexpr int R = 2;
    constexpr float kEmpty = 1e30f;  // list sentinel, above every real bound
    __shared__ __align__(16) uint16_t th[kT * DS], tl[kT * DS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lq = lane & 15, lg = lane >> 4;

    // output layout probe, as in k_knn_filter, with this instruction
    const v4f zero = {0.0f, 0.0f, 0.0f, 0.0f};
    v8bf pa, pb, pc1, pc2;
    for (int e = 0; e < 8; ++e) {
        const bool first = lg == 0 && e == 0;
        pa[e] = (__bf16)(first ? (float)lq : 0.0f);
        pb[e] = (__bf16)(first ? 1.0f : 0.0f);
        pc1[e] = (__bf16)(first ? 1.0f : 0.0f);
        pc2[e] = (__bf16)(first ? (float)lq : 0.0f);
    }
    const v4f pr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pa, pb, zero, 0, 0, 0);
    const v4f pcol = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pc1, pc2, zero, 0, 0, 0);
    int rowof[4];
    bool bad = false;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        rowof[v] = (int)pr[v] & 15;
        bad = bad || (int)pcol[v] != lq || (float)rowof[v] != pr[v];
    }
    if (bad) *layout_err = 1;

    const int64_t qb = q0 + ((int64_t)blockIdx.x * 4 + wave) * (16 * R);
    const int64_t cb = n * blockIdx.y / gridDim.y, ce = n * (blockIdx.y + 1) / gridDim.y;
    v8bf qh[R][KB], ql[R][KB];
    int32_t qi[R];
    float bd[R][kFiltL];
    int32_t bi[R][kFiltL];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int64_t q = qb + 16 * r + lq;
        const bool ok = q < q1;
        qi[r] = ok ? (int32_t)q : -1;
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
            v8u16 h = {0, 0, 0, 0, 0, 0, 0, 0}, l = h;
            if (ok) {
                h = *reinterpret_cast<const v8u16*>(bh + q * DB + 32 * kb + 8 * lg);
                l = *reinterpret_cast<const v8u16*>(bl + q * DB + 32 * kb + 8 * lg);
            }
            qh[r][kb] = __builtin_bit_cast(v8bf, h);
            ql[r][kb] = __builtin_bit_cast(v8bf, l);
        }
#pragma unroll
        for (int p = 0; p < kFiltL; ++p) { bd[r][p] = kEmpty; bi[r][p] = -1; }
    }
    // a row past the end: 2e30 in the norm slot (bf16 parts of 2e30), so its
    // bound exceeds the list sentinel and it is never kept
    const float pad = 2e30f;
    const uint16_t pad_h = bf16_rne(pad);
    const uint16_t pad_l = bf16_rne(pad - __uint_as_float((uint32_t)pad_h << 16));

    auto dots = [&](int m, v4f (&acc)[R]) {
        v8bf a_h[KB], a_l[KB];
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
            const int o = (16 * m + lq) * DS + 32 * kb + 8 * lg;
            a_h[kb] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v8u16*>(&th[o]));
            a_l[kb] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v8u16*>(&tl[o]));
        }
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = zero;
        // term-major order: consecutive MFMAs go to different query tiles, so
        // none waits for the one before it
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                acc[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_l[kb], qh[r][kb], acc[r], 0, 0, 0);
#pragma unroll
            for (int r = 0; r < R; ++r)
                acc[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h[kb], ql[r][kb], acc[r], 0, 0, 0);
#pragma unroll
            for (int r = 0; r < R; ++r)
                acc[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h[kb], qh[r][kb], acc[r], 0, 0, 0);
        }
    };
    // one test of all 16 values first: late in the scan almost no lane
    // inserts, and the per-value branches are then skipped with one
    // thr[r]: a candidate at or above it is no true neighbour (see below),
    // and none above the list's last entry can enter the list
    float thr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) thr[r] = kEmpty;
    const int pk = (k + 1 + 3) / 4 - 1;  // 4 (pk + 1) >= k + 1 entries
    auto keep = [&](int m, int64_t c0, const v4f (&acc)[R]) {
        bool any = false;
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int v = 0; v < 4; ++v) any = any || acc[r][v] < thr[r];
        if (!any) return;
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float lo = acc[r][v];
                if (lo < thr[r]) {  // thr[r] <= bd[r][kFiltL - 1] always
                    topk_insert_f<kFiltL>(bd[r], bi[r], lo, (int32_t)(c0 + 16 * m + rowof[v]));
                    thr[r] = fminf(thr[r], bd[r][kFiltL - 1]);
                }
            }
    };
    // The shared threshold, once per tile.  The four lanes of a query (l % 16
    // equal) hold lists of disjoint candidates; theta = the largest of their
    // entries pk bounds 4 (pk + 1) >= k + 1 kept bounds l, so at least k
    // candidates other than the query itself have l <= theta, hence
    // d^2 <= l + 2E <= theta + Delta (Delta = 2 (2 alpha (|q'|^2 + max |c'|^2)
    // + eabs) >= 2E) and d_k^2 <= theta + Delta.  A true neighbour has
    // l <= d^2 <= d_k^2, so one with l >= thr (theta + Delta rounded up) is not.
    auto refresh = [&]() {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float t = kEmpty;
#pragma unroll
            for (int p = 0; p < kFiltL; ++p) t = p == pk ? bd[r][p] : t;
            t = fmaxf(t, __shfl_xor(t, 16));
            t = fmaxf(t, __shfl_xor(t, 32));
            const float b = t + (qi[r] >= 0 ? fmaf(e2a, nrm[qi[r]], e2b) : 0.0f);
            const float up = b + fabsf(b) * 0x1p-20f + 1e-30f;
            thr[r] = t < kEmpty ? fminf(up, bd[r][kFiltL - 1]) : bd[r][kFiltL - 1];
        }
    };
    for (int64_t c0 = cb; c0 < ce; c0 += kT) {
        const int cnt = (int)((ce - c0) < kT ? (ce - c0) : kT);
        __syncthreads();
        for (int e = threadIdx.x; e < kT * (DB / 8); e += 256) {
            const int row = e / (DB / 8), c8 = e - row * (DB / 8);
            v8u16 h = {0, 0, 0, 0, 0, 0, 0, 0}, l = h;
            if (row < cnt) {
                h = *reinterpret_cast<const v8u16*>(ah + (c0 + row) * DB + 8 * c8);
                l = *reinterpret_cast<const v8u16*>(al + (c0 + row) * DB + 8 * c8);
            } else if (c8 == d / 8) {
                h[d % 8] = pad_h;
                l[d % 8] = pad_l;
            }
            *reinterpret_cast<v8u16*>(&th[row * DS + 8 * c8]) = h;
            *reinterpret_cast<v8u16*>(&tl[row * DS + 8 * c8]) = l;
        }
        __syncthreads();
        refresh();
        const int subs = (cnt + 15) / 16;
        v4f acc[R];
        dots(0, acc);
        for (int m = 0; m + 1 < subs; ++m) {
            v4f nxt[R];
            dots(m + 1, nxt);
            keep(m, c0, acc);
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] = nxt[r];
        }
        keep(subs - 1, c0, acc);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (qi[r] < 0) continue;
        const int64_t o =
            (((int64_t)(qi[r] - q0) * gridDim.y + blockIdx.y) * kFiltGroups + lg) * kFiltL;
#pragma unroll
        for (int p = 0; p < kFiltL; ++p) {
            kidx[o + p] = bi[r][p];
            kd2[o + p] = bi[r][p] >= 0 ? bd[r][p] : INFINITY;
        }
    }
}

// (s, j) before (bs, bj): smaller distance, then smaller index.
__device__ inline bool rank_before(double s, int32_t j, double bs, int32_t bj) {
    return s < bs || (s == bs && j < bj);
}

// Stage 2, one query per thread (see the comment above filt_tiles).  A query that
// fails the certificate, or every query with force_direct, is appended to
// qlist for k_knn.
// xp: the points as stored for k_knn (rows of dx coordinates, zero padded);
// the exact distances run over those dx coordinates in k_knn's order.
template <int DP, typename CT>
__global__ __launch_bounds__(256) void k_knn_rerank(const CT* __restrict__ xp, int dx,
                             const float* __restrict__ nrm,
                             int64_t q0, int64_t q1, int k, int lists, double alpha,
                             double eabs, const int32_t* __restrict__ kidx,
                             const float* __restrict__ kd2,
                             int force_direct, int32_t* __restrict__ oi,
                             float* __restrict__ od, int32_t* __restrict__ qlist,
                             int32_t* __restrict__ qcount) {
    const int64_t q = q0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= q1) return;
    const int slots = lists * kFiltL;
    const int64_t base = (q - q0) * slots;
    const double nq = (double)nrm[q];
    // U_k: the k-th smallest upper bound u = l + 2E over the kept candidates
    double ub[kFiltK];
#pragma unroll
    for (int p = 0; p < kFiltK; ++p) ub[p] = INFINITY;
    for (int e = 0; e < slots; ++e) {
        const int32_t j = kidx[base + e];
        if (j < 0 || j == q) continue;
        const double u = (double)kd2[base + e] + 2.0 * (2.0 * alpha * (nq + (double)nrm[j]) + eabs);
        if (!(u < ub[kFiltK - 1])) continue;
#pragma unroll
        for (int p = kFiltK - 1; p > 0; --p) {
            const bool mv = u < ub[p - 1];
            const bool here = !mv && u < ub[p];
            ub[p] = mv ? ub[p - 1] : (here ? u : ub[p]);
        }
        if (u < ub[0]) ub[0] = u;
    }
    double uk = INFINITY;
#pragma unroll
    for (int p = 0; p < kFiltK; ++p)
        if (p == k - 1) uk = ub[p];
    bool direct = force_direct != 0;
    for (int g = 0; g < lists; ++g) {
        const int64_t e = base + g * kFiltL + kFiltL - 1;
        if (kidx[e] >= 0 && !((double)kd2[e] > uk)) direct = true;
    }
    if (direct) {
        qlist[atomicAdd(qcount, 1)] = (int32_t)q;
        return;
    }
    const CT* qx = xp + q * dx;
    double bs[kFiltK];
    int32_t bj[kFiltK];
#pragma unroll
    for (int p = 0; p < kFiltK; ++p) { bs[p] = INFINITY; bj[p] = -1; }
    for (int e = 0; e < slots; ++e) {
        const int32_t j = kidx[base + e];
        if (j < 0 || j == q || (double)kd2[base + e] > uk) continue;
        double s = 0.0;
#pragma unroll
        for (int t = 0; t < DP; ++t) {
            if (t < dx) {
                const double df = (double)qx[t] - (double)xp[(int64_t)j * dx + t];
                s = fma(df, df, s);
            }
        }
        if (!rank_before(s, j, bs[kFiltK - 1], bj[kFiltK - 1])) continue;
#pragma unroll
        for (int p = kFiltK - 1; p > 0; --p) {
            const bool mv = rank_before(s, j, bs[p - 1], bj[p - 1]);
            const bool here = !mv && rank_before(s, j, bs[p], bj[p]);
            bs[p] = mv ? bs[p - 1] : (here ? s : bs[p]);
            bj[p] = mv ? bj[p - 1] : (here ? j : bj[p]);
        }
        if (rank_before(s, j, bs[0], bj[0])) { bs[0] = s; bj[0] = j; }
    }
#pragma unroll
    for (int p = 0; p < kFiltK; ++p) {
        if (p < k) {
            oi[q * k + p] = bj[p];
            od[q * k + p] = bj[p] >= 0 ? (float)sqrt(bs[p]) : INFINITY;
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

// k-NN path of later calls (testing seam): 0 automatic, 1 the direct FP64
// kernel only, 2 the filter with every query sent to the direct kernel, 3 the
// FP32 filter instead of the bf16x3 one.
std::atomic<int> g_knn_path{0};
std::atomic<int64_t> g_knn_direct{0};  // queries of the last filtered call re-run directly

// nq queries: qlist[0, nq), or 0 .. nq-1 without a list
template <int DP, int KP, typename CT>
void launch_knn_q(const CT* x, int64_t n, int k, int32_t* oi, float* od, hipStream_t s,
                  const int32_t* qlist, int64_t nq) {
    // two queries per thread halve the LDS reads per distance where the
    // registers allow it (DP <= 16, k <= 16; above, occupancy drops to one wave)
    constexpr int Q = (DP <= 16 && KP <= 16) ? 2 : 1;
    const int64_t blocks = ceil_div(nq, (int64_t)kKnnThreads * Q);
    // few queries (a small input, or the queries a filter did not certify):
    // split the candidates so that about 2048 workgroups share the scan
    const int splits =
        (int)std::max<int64_t>(1, std::min<int64_t>(16, ceil_div((int64_t)2048, blocks)));
    if (splits == 1) {
        hipLaunchKernelGGL((k_knn<DP, KP, Q, CT>), dim3(blocks), dim3(kKnnThreads), 0, s, x, n,
                           k, oi, od, qlist, nq, nullptr, nullptr);
        SI_LAUNCH_CHECK();
        return;
    }
    const int64_t slots = blocks * kKnnThreads * Q;
    DevBuf<double> pd((size_t)slots * splits * KP);
    DevBuf<int32_t> pi((size_t)slots * splits * KP);
    hipLaunchKernelGGL((k_knn<DP, KP, Q, CT>), dim3(blocks, splits), dim3(kKnnThreads), 0, s, x,
                       n, k, oi, od, qlist, nq, pd.get(), pi.get());
    hipLaunchKernelGGL(k_knn_merge, dim3(ceil_div(nq, (int64_t)256)), dim3(256), 0, s, pd.get(),
                       pi.get(), splits, KP, n, k, qlist, nq, oi, od);
    SI_LAUNCH_CHECK();
    SI_HIP(hipStreamSynchronize(s));  // pd, pi are freed on return
}

template <int DP, typename CT>
void launch_knn_k(const CT* x, int64_t n, int k, int32_t* oi, float* od, hipStream_t s,
                  const int32_t* qlist, int64_t nq) {
    if (k <= 8) launch_knn_q<DP, 8>(x, n, k, oi, od, s, qlist, nq);
    else if (k <= 16) launch_knn_q<DP, 16>(x, n, k, oi, od, s, qlist, nq);
    else if (k <= 32) launch_knn_q<DP, 32>(x, n, k, oi, od, s, qlist, nq);
    else launch_knn_q<DP, 64>(x, n, k, oi, od, s, qlist, nq);
}

// The direct kernel at the stored width dp (runtime).
template <typename CT>
void launch_knn_dp(int dp, const CT* x, int64_t n, int k, int32_t* oi, float* od, hipStream_t s,
                   const int32_t* qlist, int64_t nq) {
    switch (dp) {
        case 2: if constexpr (sizeof(CT) == 8) launch_knn_k<2>(x, n, k, oi, od, s, qlist, nq); break;
        case 4: launch_knn_k<4>(x, n, k, oi, od, s, qlist, nq); break;
        case 8: launch_knn_k<8>(x, n, k, oi, od, s, qlist, nq); break;
        case 16: launch_knn_k<16>(x, n, k, oi, od, s, qlist, nq); break;
        case 32: launch_knn_k<32>(x, n, k, oi, od, s, qlist, nq); break;
        case 48: launch_knn_k<48>(x, n, k, oi, od, s, qlist, nq); break;
        default: launch_knn_k<64>(x, n, k, oi, od, s, qlist, nq); break;
    }
}

// The two-stage path for points of dimension d (float32 or float64, stored in
// rows of dx as k_knn reads them), filtered at DP = 4 KS coordinates
// coordinates (k <= kFiltK).  Stage 1 runs on the bf16 matrix cores (three
// bf16 products per FP32 product) when the squared norms of the centred
// points are at most 1e20 and d + 2 <= 96; if that leaves more than 1 % of the
// queries uncertified, or otherwise, the FP32 matrix cores run it.  Returns
// false, having written nothing, when a squared norm is not finite or above
// 1e36; the caller then runs the direct kernel.
template <int KS, typename CT>
bool knn_filtered(const CT* xp, int dx, int64_t n, int d, int k, int32_t* oi, float* od,
                  hipStream_t s) {
    constexpr int DP = 4 * KS;
    DevBuf<double> part((size_t)kColParts * DP), mean(DP);
    // means over every stored column (the padding's are 0)
    hipLaunchKernelGGL((k_col_partials<CT>), dim3(kColParts), dim3(64), 0, s, xp, n, dx, dx,
                       part.get());
    hipLaunchKernelGGL(k_col_mean, dim3(1), dim3(1024), 0, s, part.get(), kColParts, n, dx,
                       mean.get());
    SI_LAUNCH_CHECK();
    DevBuf<float> xc((size_t)n * DP), nrm(n);
    DevBuf<int32_t> ctl(4);  // norm max bits, layout error, direct-query count
    SI_HIP(hipMemsetAsync(ctl.get(), 0, 4 * sizeof(int32_t), s));
    hipLaunchKernelGGL((k_center_norms<DP, CT>), dim3(ceil_div(n, 256)), dim3(256), 0, s, xp, n,
                       dx, d, mean.get(), xc.get(), nrm.get(),
                       reinterpret_cast<unsigned*>(ctl.get()));
    SI_LAUNCH_CHECK();
    uint32_t mbits = 0;
    SI_HIP(hipMemcpyAsync(&mbits, ctl.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    SI_HIP(hipStreamSynchronize(s));
    float m2 = 0.0f;
    std::memcpy(&m2, &mbits, sizeof(float));
    if (!(m2 <= 1e36f)) return false;  // also false for NaN
    const int mode = g_knn_path.load();
    const int kb = (d + 2 + 31) / 32;
    const bool bf16x3 = mode != 3 && m2 <= 1e20f && kb <= 3;
    const double u = std::ldexp(1.0, -24);
    // FP32 stage 1: |a - d^2| <= (DP + 12) u (|q'|^2 + |c'|^2) + eabs from the
    // FP32 rounding of the centred coordinates, the norms, the MFMA sums and
    // the epilogue (with (x + y)^2 <= 2 (x^2 + y^2)); 2 alpha = 2 (DP + 16) u
    // leaves room for the rounding of l itself.  eabs covers flushed subnormal
    // products.
    const double alpha32 = (double)(DP + 16) * u;
    // bf16x3 stage 1: the dropped part products (3.02 2^-16 |x y| per
    // coordinate, 2^-16 of each norm slot), the FP32 sums of 3 DB terms and
    // the rounding of the centred coordinates and norms give
    // (4.02 2^-16 + (6 DB + 10) 2^-24) (|q'|^2 + |c'|^2); 2 alpha doubles it.
    const int db = 32 * kb;
    const double alpha3 = 4.02 * std::ldexp(1.0, -16) + (6.0 * db + 10.0) * u;
    const double eabs = 96.0 * 1e-30;
    const int qpb = bf16x3 ? 4 * 16 * 2 : filt_queries<KS>();
    // candidate splits: about 2048 workgroups, at most 16 splits
    const int64_t qblocks = ceil_div(n, (int64_t)qpb);
    const int splits =
        (int)std::max<int64_t>(1, std::min<int64_t>(16, ceil_div((int64_t)2048, qblocks)));
    const int lists = kFiltGroups * splits;
    // query chunks keep the kept lists within 4 GiB
    const int64_t chunk = std::min<int64_t>(n, ((int64_t)1 << 29) / (lists * kFiltL));
    DevBuf<int32_t> kidx((size_t)chunk * lists * kFiltL), qlist(n);
    DevBuf<float> kd2((size_t)chunk * lists * kFiltL);
    DevBuf<uint16_t> ah, al, bh, bl;
    if (bf16x3) {
        const size_t rows = (size_t)n * db;
        ah.alloc(rows); al.alloc(rows); bh.alloc(rows); bl.alloc(rows);
        hipLaunchKernelGGL(k_split_rows, dim3(ceil_div((int64_t)rows, (int64_t)256)), dim3(256), 0,
                           s, xc.get(), nrm.get(), n, d, DP, db, (float)(1.0 - 2.0 * alpha3),
                           (float)eabs, ah.get(), al.get(), bh.get(), bl.get());
        SI_LAUNCH_CHECK();
    }
    // Delta of the bf16 filter's shared threshold: 2 (2 alpha (nq + max nc) +
    // eabs) = e2a nq + e2b, rounded up
    const float e2a = (float)(4.0 * alpha3 * (1.0 + 1e-6));
    const float e2b = (float)((4.0 * alpha3 * (double)m2 + 2.0 * eabs) * (1.0 + 1e-6));
    const float e2a32 = (float)(4.0 * alpha32 * (1.0 + 1e-6));
    const float e2b32 = (float)((4.0 * alpha32 * (double)m2 + 2.0 * eabs) * (1.0 + 1e-6));
    const int force = mode == 2 ? 1 : 0;
    auto pass = [&](bool three) -> int32_t {
        SI_HIP(hipMemsetAsync(ctl.get() + 2, 0, sizeof(int32_t), s));
        const double alpha = three ? alpha3 : alpha32;
        for (int64_t q0 = 0; q0 < n; q0 += chunk) {
            const int64_t q1 = std::min(n, q0 + chunk);
            // queries per workgroup of the kernel this pass runs
            const int64_t qw = three ? qpb : filt_queries<KS>();
            const dim3 grid((unsigned)ceil_div(q1 - q0, qw), (unsigned)splits);
            if (three) {
                switch (kb) {
                    case 1:
                        hipLaunchKernelGGL((k_knn_filter3<1>), grid, dim3(256), 0, s, ah.get(),
                                           al.get(), bh.get(), bl.get(), n, d, k, nrm.get(), e2a,
                                           e2b, q0, q1, kidx.get(), kd2.get(), ctl.get() + 1);
                        break;
                    case 2:
                        hipLaunchKernelGGL((k_knn_filter3<2>), grid, dim3(256), 0, s, ah.get(),
                                           al.get(), bh.get(), bl.get(), n, d, k, nrm.get(), e2a,
                                           e2b, q0, q1, kidx.get(), kd2.get(), ctl.get() + 1);
                        break;
                    default:
                        hipLaunchKernelGGL((k_knn_filter3<3>), grid, dim3(256), 0, s, ah.get(),
                                           al.get(), bh.get(), bl.get(), n, d, k, nrm.get(), e2a,
                                           e2b, q0, q1, kidx.get(), kd2.get(), ctl.get() + 1);
                        break;
                }
            } else {
                hipLaunchKernelGGL((k_knn_filter<KS>), grid, dim3(256), 0, s, xc.get(), nrm.get(),
                                   n, q0, q1, (float)(1.0 - 2.0 * alpha), (float)eabs, k,
                                   e2a32, e2b32, kidx.get(), kd2.get(), ctl.get() + 1);
            }
            hipLaunchKernelGGL((k_knn_rerank<DP, CT>), dim3(ceil_div(q1 - q0, 256)), dim3(256), 0,
                               s, xp, dx, nrm.get(), q0, q1, k, lists, alpha, eabs, kidx.get(),
                               kd2.get(), force, oi, od, qlist.get(), ctl.get() + 2);
            SI_LAUNCH_CHECK();
        }
        int32_t h[2] = {0, 0};
        SI_HIP(hipMemcpyAsync(h, ctl.get() + 1, sizeof(h), hipMemcpyDeviceToHost, s));
        SI_HIP(hipStreamSynchronize(s));
        SI_CHECK(h[0] == 0, kErrInternal, "knn: unexpected MFMA output layout");
        return h[1];
    };
    int32_t direct = pass(bf16x3);
    if (bf16x3 && !force && direct > n / 100) direct = pass(false);
    g_knn_direct.store(direct);
    if (direct > 0) launch_knn_dp<CT>(dx, xp, n, k, oi, od, s, qlist.get(), direct);
    SI_HIP(hipStreamSynchronize(s));  // the buffers are freed on return
    return true;
}

// dp: the filter's width (a multiple of 4 >= d); dx: the stored width of xp
template <typename CT>
bool knn_filtered_dp(int dp, const CT* xp, int dx, int64_t n, int d, int k, int32_t* oi,
                     float* od, hipStream_t s) {
    switch (dp) {
        case 4: return knn_filtered<1>(xp, dx, n, d, k, oi, od, s);
        case 8: return knn_filtered<2>(xp, dx, n, d, k, oi, od, s);
        case 16: return knn_filtered<4>(xp, dx, n, d, k, oi, od, s);
        case 32: return knn_filtered<8>(xp, dx, n, d, k, oi, od, s);
        case 48: return knn_filtered<12>(xp, dx, n, d, k, oi, od, s);
        default: return knn_filtered<16>(xp, dx, n, d, k, oi, od, s);
    }
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
    if (g_knn_path.load() != 1 && k <= kFiltK &&
        knn_filtered_dp<CT>(padded_dim(d, 4), xp.get(), dp, n, (int)d, (int)k, idx, dist, s))
        return;
    launch_knn_dp<CT>(dp, xp.get(), n, (int)k, idx, dist, s, nullptr, n);
    SI_HIP(hipStreamSynchronize(s));  // xp is freed on return
}

}  // namespace

void knn_set_path(int mode) { g_knn_path.store(mode); }
int64_t knn_direct_count() { return g_knn_direct.load(); }

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
