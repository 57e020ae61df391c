// k_block.hip -- tall-skinny block kernels of the block eigensolvers:
// FP64-accumulated Gram matrices on the FP64 matrix cores, block x small
// matrix products (the fused AXPY updates), residual blocks, Philox normal
// fills and conversions.
//
// Gram G = A^T B, A [n x p], B [n x q], p, q <= 64.  Each wavefront walks
// 4 rows per step and feeds v_mfma_f64_16x16x4f64 with A^T (16 x 4) and
// B (4 x 16) tiles; FP32 inputs are widened to FP64 in registers, so every
// product and every sum is FP64.  The four wavefronts of a workgroup reduce in
// LDS and write one partial per workgroup; a second kernel sums the partials
// in a fixed order (deterministic).
#include "si_device.hpp"

namespace si {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kGramBlocks = 1024;

// B may be split in two column blocks: columns c < q1 come from b, columns
// q1 <= c < q from b2 (column c - q1), so [Y^T Y | Y^T LY] is one pass.
template <typename T, int PT, int QT>
__global__ __launch_bounds__(256) void k_gram(int64_t n, const T* __restrict__ a, int lda, int p,
                                              const T* __restrict__ b, int ldb, int q1,
                                              const T* __restrict__ b2, int ldb2, int q,
                                              double* __restrict__ partial,
                                              int64_t rows_per_block,
                                              const int64_t* __restrict__ blk_rows,
                                              const int32_t* __restrict__ gate) {
    if (gate != nullptr && *gate == 0) return;
    __shared__ double red[PT * 16][QT * 16 + 1];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int li = lane & 15;
    const int lk = lane >> 4;
    v4d acc[PT][QT];
#pragma unroll
    for (int ti = 0; ti < PT; ++ti)
#pragma unroll
        for (int tj = 0; tj < QT; ++tj) acc[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};
    int64_t r0, r1;
    if (blk_rows) {  // segmented: rows [r0, r1) of one segment
        r0 = blk_rows[2 * blockIdx.x];
        r1 = blk_rows[2 * blockIdx.x + 1];
    } else {
        r0 = (int64_t)blockIdx.x * rows_per_block;
        r1 = (r0 + rows_per_block) < n ? (r0 + rows_per_block) : n;
    }
    // the first B block equal to A: read it once
    const bool same = (a == b) && (lda == ldb) && (q1 == p) && (q1 % 16 == 0 || q1 == q);
    // UR row groups of 4 rows per wave in flight: all loads first, then MFMAs
    constexpr int UR = 4;
    for (int64_t rb = r0 + wave * 4; rb < r1; rb += 16 * UR) {
        double av[UR][PT], bv[UR][QT];
#pragma unroll
        for (int g = 0; g < UR; ++g) {
            const int64_t row = rb + 16 * g + lk;
            const bool ok = row < r1;
#pragma unroll
            for (int t = 0; t < PT; ++t) {
                const int c = t * 16 + li;
                av[g][t] = (ok && c < p) ? (double)a[row * lda + c] : 0.0;
            }
#pragma unroll
            for (int t = 0; t < QT; ++t) {
                const int c = t * 16 + li;
                if (same && c < q1 && t < PT) {
                    bv[g][t] = av[g][t < PT ? t : 0];
                } else if (c < q1) {
                    bv[g][t] = ok ? (double)b[row * ldb + c] : 0.0;
                } else {
                    bv[g][t] = (ok && c < q) ? (double)b2[row * ldb2 + (c - q1)] : 0.0;
                }
            }
        }
#pragma unroll
        for (int g = 0; g < UR; ++g)
#pragma unroll
            for (int ti = 0; ti < PT; ++ti)
#pragma unroll
                for (int tj = 0; tj < QT; ++tj)
                    acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[g][ti], bv[g][tj],
                                                                       acc[ti][tj], 0, 0, 0);
    }
    // D[i][j] of a tile sits in lane (j + 16 * (i % 4)), element i / 4
    // (measured on gfx950 with an integer-valued probe).
    for (int w = 0; w < 4; ++w) {
        if (wave == w) {
#pragma unroll
            for (int ti = 0; ti < PT; ++ti)
#pragma unroll
                for (int tj = 0; tj < QT; ++tj)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int i = ti * 16 + lk + 4 * v;
                        const int j = tj * 16 + li;
                        if (w == 0) red[i][j] = acc[ti][tj][v];
                        else red[i][j] += acc[ti][tj][v];
                    }
        }
        __syncthreads();
    }
    double* out = partial + (int64_t)blockIdx.x * p * q;
    for (int e = threadIdx.x; e < p * q; e += 256) out[e] = red[e / q][e % q];
}

// Narrow Gram (p, q <= 4): one thread per row (grid stride) keeps the p x q
// running sums in registers (products and sums in FP64); a workgroup reduces
// them with wave butterflies and a fixed-order sum of its 4 waves into one
// partial, laid out as k_gram's (p x q row-major per workgroup).
template <typename T>
__global__ __launch_bounds__(256) void k_gram_narrow(int64_t n, const T* __restrict__ a, int lda,
                                                     int p, const T* __restrict__ b, int ldb,
                                                     int q1, const T* __restrict__ b2, int ldb2,
                                                     int q, double* __restrict__ partial) {
    double acc[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256) {
        double av[4], bv[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            av[c] = c < p ? (double)a[i * lda + c] : 0.0;
            bv[c] = c < q1 ? (double)b[i * ldb + c]
                           : (c < q ? (double)b2[i * ldb2 + (c - q1)] : 0.0);
        }
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) acc[x * 4 + y] += av[x] * bv[y];
    }
    __shared__ double red[4][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[e] += __shfl_xor(acc[e], off, 64);
    if (lane == 0)
#pragma unroll
        for (int e = 0; e < 16; ++e) red[wave][e] = acc[e];
    __syncthreads();
    const int t = threadIdx.x;
    if (t < p * q) {
        const int x = t / q, y = t % q;
        const int e = x * 4 + y;
        partial[(int64_t)blockIdx.x * p * q + t] =
            ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
    }
}

// Sum the per-workgroup partials, g[e] = sum_b partial[b][e], in a fixed order
// (deterministic).  A workgroup owns 8 outputs; 32 groups of 8 lanes stride the
// partials, each lane keeping 8 independent loads in flight, and the group sums
// are added by a fixed tree in LDS.
__global__ __launch_bounds__(256) void k_reduce_partials(const double* __restrict__ partial,
                                                         int nparts, int pq,
                                                         double* __restrict__ g) {
    constexpr int E = 8, G = 256 / E, D = 8;
    __shared__ double red[G][E + 1];
    const int lane = threadIdx.x % E, grp = threadIdx.x / E;
    const int e = blockIdx.x * E + lane;
    double acc[D];
#pragma unroll
    for (int u = 0; u < D; ++u) acc[u] = 0.0;
    if (e < pq) {
        for (int b = grp; b < nparts; b += D * G) {
#pragma unroll
            for (int u = 0; u < D; ++u) {
                const int bb = b + u * G;
                if (bb < nparts) acc[u] += partial[(int64_t)bb * pq + e];
            }
        }
    }
    double t = 0.0;
#pragma unroll
    for (int u = 0; u < D; ++u) t += acc[u];
    red[grp][lane] = t;
    __syncthreads();
    for (int h = G / 2; h > 0; h >>= 1) {
        if (grp < h) red[grp][lane] += red[grp + h][lane];
        __syncthreads();
    }
    if (grp == 0 && e < pq) g[e] = red[0][lane];
}

// Segmented: g[c][e] = sum of partial[b][e] over b in [seg_blk[c], seg_blk[c+1]),
// segment c = blockIdx.y; same fixed order as above.
__global__ __launch_bounds__(256) void k_reduce_partials_seg(const double* __restrict__ partial,
                                                             const int32_t* __restrict__ seg_blk,
                                                             int pq, double* __restrict__ g,
                                                             const int32_t* __restrict__ gate) {
    if (gate != nullptr && *gate == 0) return;
    constexpr int E = 8, G = 256 / E, D = 8;
    __shared__ double red[G][E + 1];
    const int lane = threadIdx.x % E, grp = threadIdx.x / E;
    const int e = blockIdx.x * E + lane;
    const int b0 = seg_blk[blockIdx.y], b1 = seg_blk[blockIdx.y + 1];
    double acc[D];
#pragma unroll
    for (int u = 0; u < D; ++u) acc[u] = 0.0;
    if (e < pq) {
        for (int b = b0 + grp; b < b1; b += D * G) {
#pragma unroll
            for (int u = 0; u < D; ++u) {
                const int bb = b + u * G;
                if (bb < b1) acc[u] += partial[(int64_t)bb * pq + e];
            }
        }
    }
    double t = 0.0;
#pragma unroll
    for (int u = 0; u < D; ++u) t += acc[u];
    red[grp][lane] = t;
    __syncthreads();
    for (int h = G / 2; h > 0; h >>= 1) {
        if (grp < h) red[grp][lane] += red[grp + h][lane];
        __syncthreads();
    }
    if (grp == 0 && e < pq) g[(int64_t)blockIdx.y * pq + e] = red[0][lane];
}

void reduce_partials_seg(const double* partial, const int32_t* seg_blk, int nseg, int pq,
                         double* g, hipStream_t s, const int32_t* gate = nullptr) {
    hipLaunchKernelGGL(k_reduce_partials_seg, dim3(ceil_div(pq, 8), nseg), dim3(256), 0, s,
                       partial, seg_blk, pq, g, gate);
    SI_LAUNCH_CHECK();
}

void reduce_partials(const double* partial, int nparts, int pq, double* g, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(ceil_div(pq, 8)), dim3(256), 0, s, partial, nparts,
                       pq, g);
    SI_LAUNCH_CHECK();
}

template <typename T>
void gram_impl(int64_t n, const T* a, int lda, int p, const T* b, int ldb, int q1, const T* b2,
               int ldb2, int q, double* g, double* ws, hipStream_t s,
               const SegWork* seg = nullptr, const int32_t* gate = nullptr) {
    SI_CHECK(p >= 1 && q >= 1 && p <= 64 && q <= 64 && q1 >= 1 && q1 <= q, kErrInvalidArgument,
             "gram: width out of range");
    SI_CHECK(q1 == q || b2 != nullptr, kErrInvalidArgument, "gram: second block missing");
    int64_t nblocks = 1, rpb = 16;
    const int64_t* blk_rows = nullptr;
    if (seg) {  // the work list gives every block's rows
        nblocks = seg->nblk;
        blk_rows = seg->blk_rows;
        SI_CHECK(nblocks >= 1, kErrInvalidArgument, "gram: empty segment work list");
    } else if (n > 0) {
        nblocks = ceil_div(n, 64);
        if (nblocks > kGramBlocks) nblocks = kGramBlocks;
        rpb = ceil_div(ceil_div(n, nblocks), 16) * 16;
        nblocks = ceil_div(n, rpb);
    }
    if (!seg && p <= 4 && q <= 4) {  // narrow blocks: no 16-wide MFMA tiles of padding
        hipLaunchKernelGGL((k_gram_narrow<T>), dim3(nblocks), dim3(256), 0, s, n, a, lda, p, b,
                           ldb, q1, b2, ldb2, q, ws);
        SI_LAUNCH_CHECK();
        reduce_partials(ws, (int)nblocks, p * q, g, s);
        return;
    }
    const int pt = (p + 15) / 16, qt = (q + 15) / 16;
#define SI_GRAM_CASE(PT_, QT_)                                                                   \
    if (pt == PT_ && qt == QT_) {                                                                \
        hipLaunchKernelGGL((k_gram<T, PT_, QT_>), dim3(nblocks), dim3(256), 0, s, n, a, lda, p, b, \
                           ldb, q1, b2, ldb2, q, ws, rpb, blk_rows, gate);                        \
    } else
    SI_GRAM_CASE(1, 1) SI_GRAM_CASE(1, 2) SI_GRAM_CASE(1, 3) SI_GRAM_CASE(1, 4)
    SI_GRAM_CASE(2, 1) SI_GRAM_CASE(2, 2) SI_GRAM_CASE(2, 3) SI_GRAM_CASE(2, 4)
    SI_GRAM_CASE(3, 1) SI_GRAM_CASE(3, 2) SI_GRAM_CASE(3, 3) SI_GRAM_CASE(3, 4)
    SI_GRAM_CASE(4, 1) SI_GRAM_CASE(4, 2) SI_GRAM_CASE(4, 3) SI_GRAM_CASE(4, 4)
    { throw Error(kErrInternal, "gram: no kernel for this shape"); }
#undef SI_GRAM_CASE
    SI_LAUNCH_CHECK();
    if (seg)
        reduce_partials_seg(ws, seg->seg_blk, seg->nseg, p * q, g, s, gate);
    else
        reduce_partials(ws, (int)nblocks, p * q, g, s);
}

// ------------------------------------------------------------------ block products
// Out[:, 0:q] = A[:, 0:p] C (+ B[:, 0:p2] D).  A workgroup owns R rows; their
// A (then B) tile is staged through LDS with coalesced loads, the small matrix
// (C, then D) sits in LDS in the data precision, and TPR threads per row each
// produce CPT contiguous output columns (coalesced stores).  Rows are only
// written after both tiles were read, so out may alias a (in place).
template <typename T, int QMAX>
__global__ __launch_bounds__(256) void k_block_mul(int64_t n, const T* __restrict__ a, int lda,
                                                   int p, const double* __restrict__ c, int q,
                                                   const T* __restrict__ b, int ldb, int p2,
                                                   const double* __restrict__ d, T* out, int ldo) {
    constexpr int R = sizeof(T) == 8 ? 32 : 64;
    constexpr int TPR = 256 / R;
    constexpr int CPT = (QMAX + TPR - 1) / TPR;
    constexpr int LDT = 65;
    __shared__ T cm[64 * QMAX];
    __shared__ T tile[R * LDT];
    const int t = threadIdx.x;
    const int r = t / TPR, g = t % TPR;
    const int64_t row0 = (int64_t)blockIdx.x * R;
    const int rows = (n - row0) < R ? (int)(n - row0) : R;
    T acc[CPT];
#pragma unroll
    for (int u = 0; u < CPT; ++u) acc[u] = T(0);
    for (int term = 0; term < 2; ++term) {
        const T* src = term == 0 ? a : b;
        if (!src) break;
        const int ld = term == 0 ? lda : ldb;
        const int pp = term == 0 ? p : p2;
        const double* cc = term == 0 ? c : d;
        __syncthreads();
        for (int e = t; e < pp * q; e += 256) cm[(e / q) * QMAX + e % q] = (T)cc[e];
        for (int e = t; e < rows * pp; e += 256) {
            const int rr = e / pp, kk = e % pp;
            tile[rr * LDT + kk] = src[(row0 + rr) * ld + kk];
        }
        __syncthreads();
        if (r < rows) {
            for (int k = 0; k < pp; ++k) {
                const T av = tile[r * LDT + k];
#pragma unroll
                for (int u = 0; u < CPT; ++u) {
                    const int j = g * CPT + u;
                    if (j < QMAX) acc[u] += av * cm[k * QMAX + j];
                }
            }
        }
    }
    if (r < rows) {
        T* o = out + (row0 + r) * ldo;
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int j = g * CPT + u;
            if (j < q) o[j] = acc[u];
        }
    }
}

// Narrow block product (p, p2, q <= 4): one thread per row (grid stride), the
// small matrices in registers in the data precision, the same term and k order
// as k_block_mul.  A thread reads its row before writing it: out may alias a.
template <typename T>
__global__ __launch_bounds__(256) void k_block_mul_narrow(int64_t n, const T* a, int lda, int p,
                                                          const double* __restrict__ c, int q,
                                                          const T* __restrict__ b, int ldb, int p2,
                                                          const double* __restrict__ d, T* out,
                                                          int ldo) {
    T cm[16], dm[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int k = e / 4, j = e % 4;
        cm[e] = (k < p && j < q) ? (T)c[k * q + j] : T(0);
        dm[e] = (b && k < p2 && j < q) ? (T)d[k * q + j] : T(0);
    }
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256) {
        T av[4], bv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            av[k] = k < p ? a[i * lda + k] : T(0);
            bv[k] = (b && k < p2) ? b[i * ldb + k] : T(0);
        }
        T acc[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[j] += av[k] * cm[k * 4 + j];
        if (b) {
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[j] += bv[k] * dm[k * 4 + j];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (j < q) out[i * ldo + j] = acc[j];
    }
}

// Out[n x 16] = A[n x 16] C[16 x 16] in FP64 on v_mfma_f64_16x16x4f64.  Each
// wavefront owns 16-row tiles (grid stride).  Lane (li = lane % 16, lk =
// lane / 16) feeds A[r0 + li][4s + lk] and C[4s + lk][li] for the k-steps s =
// 0..3; D[i][j] comes back in lane j + 16 (i % 4), element i / 4, so each
// store instruction writes 4 whole 128-byte rows.  A wavefront reads all of
// its tile before it writes it, and no other wavefront touches those rows:
// the product may overwrite A (out == a).
__global__ __launch_bounds__(256) void k_block_mul16_mfma(int64_t n, const double* a, int lda,
                                                          const double* __restrict__ c,
                                                          double* out, int ldo) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    double cb[4];
#pragma unroll
    for (int st = 0; st < 4; ++st) cb[st] = c[(4 * st + lk) * 16 + li];
    const int64_t tiles = (n + 15) / 16;
    for (int64_t t = (int64_t)blockIdx.x * 4 + wave; t < tiles; t += (int64_t)gridDim.x * 4) {
        const int64_t r0 = t * 16;
        const int64_t ra = r0 + li;
        double av[4];
#pragma unroll
        for (int st = 0; st < 4; ++st) av[st] = ra < n ? a[ra * lda + 4 * st + lk] : 0.0;
        v4d acc = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int st = 0; st < 4; ++st)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[st], cb[st], acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t row = r0 + lk + 4 * v;
            if (row < n) out[row * ldo + li] = acc[v];
        }
    }
}

template <typename T>
void block_mul_impl(int64_t n, const T* a, int lda, int p, const double* c, int q, const T* b,
                    int ldb, int p2, const double* d, T* out, int ldo, hipStream_t s) {
    SI_CHECK(p <= 64 && q <= 64 && p2 <= 64, kErrInvalidArgument, "block_mul: width > 64");
    SI_CHECK(out != b || b == nullptr, kErrInvalidArgument, "block_mul: out aliases the second term");
    if constexpr (sizeof(T) == 8) {
        if (p == 16 && q == 16 && b == nullptr) {  // the 16-wide TS-GEMM on the FP64 matrix cores
            int64_t blocks = ceil_div(ceil_div(n, (int64_t)16), (int64_t)4);
            if (blocks > 2048) blocks = 2048;
            if (blocks < 1) blocks = 1;
            hipLaunchKernelGGL(k_block_mul16_mfma, dim3(blocks), dim3(256), 0, s, n,
                               (const double*)a, lda, c, (double*)out, ldo);
            SI_LAUNCH_CHECK();
            return;
        }
    }
    if (p <= 4 && p2 <= 4 && q <= 4) {  // narrow blocks: a thread per row, no LDS staging
        int64_t nb = ceil_div(n, (int64_t)256);
        if (nb > 8192) nb = 8192;
        if (nb < 1) nb = 1;
        hipLaunchKernelGGL((k_block_mul_narrow<T>), dim3(nb), dim3(256), 0, s, n, a, lda, p, c, q,
                           b, ldb, p2, d, out, ldo);
        SI_LAUNCH_CHECK();
        return;
    }
    const int64_t blocks = ceil_div(n, sizeof(T) == 8 ? 32 : 64);
    if (q <= 4)
        hipLaunchKernelGGL((k_block_mul<T, 4>), dim3(blocks), dim3(256), 0, s, n, a, lda, p, c, q, b, ldb, p2, d, out, ldo);
    else if (q <= 8)
        hipLaunchKernelGGL((k_block_mul<T, 8>), dim3(blocks), dim3(256), 0, s, n, a, lda, p, c, q, b, ldb, p2, d, out, ldo);
    else if (q <= 16)
        hipLaunchKernelGGL((k_block_mul<T, 16>), dim3(blocks), dim3(256), 0, s, n, a, lda, p, c, q, b, ldb, p2, d, out, ldo);
    else if (q <= 32)
        hipLaunchKernelGGL((k_block_mul<T, 32>), dim3(blocks), dim3(256), 0, s, n, a, lda, p, c, q, b, ldb, p2, d, out, ldo);
    else
        hipLaunchKernelGGL((k_block_mul<T, 64>), dim3(blocks), dim3(256), 0, s, n, a, lda, p, c, q, b, ldb, p2, d, out, ldo);
    SI_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ fused RR update
// x = y C, ax = ly C (same C), and for the first k columns the partial sums
// ||ax_j - theta_j x_j||^2 and ||x_j||^2 (partial[block][j], partial[block][k + j]).
// A workgroup walks R-row tiles with a grid stride (at most kRrBlocks
// workgroups, so few partials); TPR threads per row own CPT contiguous columns
// each and accumulate the norms of the wanted columns they own.
constexpr int kRrBlocks = 2048;

// One R-row tile of the update; accumulates the wanted-column norms of the
// tile's rows into rr2/xx2 (per owned column).  cm and tile are LDS arrays.
template <typename T, int QMAX>
__device__ inline void rr_tile(int64_t row0, int rows, const T* __restrict__ y,
                               const T* __restrict__ ly, int ld, int m, int k, const T* cm,
                               T* tile, const double* th, T* __restrict__ x, T* __restrict__ ax,
                               double* rr2, double* xx2) {
    constexpr int R = sizeof(T) == 8 ? 32 : 64;
    constexpr int TPR = 256 / R;
    constexpr int CPT = QMAX / TPR;
    constexpr int LDT = 65;
    const int t = threadIdx.x;
    const int r = t / TPR, g = t % TPR;
    T accx[CPT], acca[CPT];
#pragma unroll
    for (int u = 0; u < CPT; ++u) { accx[u] = T(0); acca[u] = T(0); }
    for (int pass = 0; pass < 2; ++pass) {
        const T* src = pass == 0 ? y : ly;
        __syncthreads();
        for (int e = t; e < rows * m; e += 256) {
            const int rr = e / m, kk = e % m;
            tile[rr * LDT + kk] = src[(row0 + rr) * ld + kk];
        }
        __syncthreads();
        if (r < rows) {
            for (int kk = 0; kk < m; ++kk) {
                const T av = tile[r * LDT + kk];
#pragma unroll
                for (int u = 0; u < CPT; ++u) {
                    const T cv = cm[kk * QMAX + g * CPT + u];
                    if (pass == 0) accx[u] += av * cv;
                    else acca[u] += av * cv;
                }
            }
        }
    }
    if (r < rows) {
        T* xo = x + (row0 + r) * ld;
        T* ao = ax + (row0 + r) * ld;
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int j = g * CPT + u;
            if (j < m) {
                xo[j] = accx[u];
                ao[j] = acca[u];
                if (j < k) {
                    const double xv = (double)accx[u];
                    const double rv = (double)acca[u] - th[u] * xv;
                    rr2[u] += rv * rv;
                    xx2[u] += xv * xv;
                }
            }
        }
    }
}

// Block-level fixed-order sum of the owned wanted-column norms into
// partial[pidx][0..2k).
template <typename T, int QMAX>
__device__ inline void rr_norms_out(int k, const double* rr2, const double* xx2,
                                    double (*red)[2 * 8 + 1], double* __restrict__ partial,
                                    int64_t pidx) {
    constexpr int R = sizeof(T) == 8 ? 32 : 64;
    constexpr int TPR = 256 / R;
    constexpr int CPT = QMAX / TPR;
    const int t = threadIdx.x;
    const int r = t / TPR, g = t % TPR;
    __syncthreads();
    // every (row slot, wanted column) has exactly one owner thread
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        const int j = g * CPT + u;
        if (j < k) { red[r][j] = rr2[u]; red[r][8 + j] = xx2[u]; }
    }
    __syncthreads();
    if (t < 2 * k) {
        const int j = t < k ? t : 8 + (t - k);
        double sacc = 0.0;
        for (int q = 0; q < R; ++q) sacc += red[q][j];
        partial[pidx * 2 * k + t] = sacc;
    }
}

template <typename T, int QMAX>
__global__ __launch_bounds__(256) void k_rr_update(int64_t n, const T* __restrict__ y,
                                                   const T* __restrict__ ly, int ld, int m,
                                                   const double* __restrict__ c,
                                                   const double* __restrict__ theta, int k,
                                                   T* __restrict__ x, T* __restrict__ ax,
                                                   double* __restrict__ partial) {
    constexpr int R = sizeof(T) == 8 ? 32 : 64;
    constexpr int TPR = 256 / R;
    constexpr int CPT = QMAX / TPR;
    static_assert(CPT * TPR == QMAX, "rr_update: QMAX must be a multiple of TPR");
    constexpr int LDT = 65;
    __shared__ T cm[64 * QMAX];
    __shared__ T tile[R * LDT];
    __shared__ double red[R][2 * 8 + 1];
    const int t = threadIdx.x;
    const int g = t % TPR;
    for (int e = t; e < m * m; e += 256) cm[(e / m) * QMAX + e % m] = (T)c[e];
    double rr2[CPT], xx2[CPT], th[CPT];
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        const int j = g * CPT + u;
        rr2[u] = 0.0;
        xx2[u] = 0.0;
        th[u] = j < k ? theta[j] : 0.0;
    }
    const int64_t tiles = (n + R - 1) / R;
    for (int64_t tl = blockIdx.x; tl < tiles; tl += gridDim.x) {
        const int64_t row0 = tl * R;
        const int rows = (n - row0) < R ? (int)(n - row0) : R;
        rr_tile<T, QMAX>(row0, rows, y, ly, ld, m, k, cm, tile, th, x, ax, rr2, xx2);
    }
    rr_norms_out<T, QMAX>(k, rr2, xx2, red, partial, blockIdx.x);
}

// Segmented: workgroup b updates tile b (rows blk_rows[2b] .. blk_rows[2b+1],
// at most R, inside segment blk_seg[b]) with that segment's C (c + seg*m*m)
// and Ritz values (theta + seg*m); one partial per tile.
template <typename T, int QMAX>
__global__ __launch_bounds__(256) void k_rr_update_seg(const int64_t* __restrict__ blk_rows,
                                                       const int32_t* __restrict__ blk_seg,
                                                       const T* __restrict__ y,
                                                       const T* __restrict__ ly, int ld, int m,
                                                       const double* __restrict__ c,
                                                       const double* __restrict__ theta, int k,
                                                       T* __restrict__ x, T* __restrict__ ax,
                                                       double* __restrict__ partial,
                                                       const int32_t* __restrict__ gate) {
    constexpr int R = sizeof(T) == 8 ? 32 : 64;
    constexpr int TPR = 256 / R;
    constexpr int CPT = QMAX / TPR;
    static_assert(CPT * TPR == QMAX, "rr_update: QMAX must be a multiple of TPR");
    constexpr int LDT = 65;
    __shared__ T cm[64 * QMAX];
    __shared__ T tile[R * LDT];
    __shared__ double red[R][2 * 8 + 1];
    if (gate != nullptr && *gate == 0) return;
    const int t = threadIdx.x;
    const int g = t % TPR;
    const int sg = blk_seg[blockIdx.x];
    const double* cs = c + (int64_t)sg * m * m;
    const double* ts = theta + (int64_t)sg * m;
    for (int e = t; e < m * m; e += 256) cm[(e / m) * QMAX + e % m] = (T)cs[e];
    double rr2[CPT], xx2[CPT], th[CPT];
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        const int j = g * CPT + u;
        rr2[u] = 0.0;
        xx2[u] = 0.0;
        th[u] = j < k ? ts[j] : 0.0;
    }
    const int64_t row0 = blk_rows[2 * blockIdx.x];
    const int rows = (int)(blk_rows[2 * blockIdx.x + 1] - row0);
    rr_tile<T, QMAX>(row0, rows, y, ly, ld, m, k, cm, tile, th, x, ax, rr2, xx2);
    rr_norms_out<T, QMAX>(k, rr2, xx2, red, partial, blockIdx.x);
}

template <typename T>
void rr_update_impl(int64_t n, const T* y, const T* ly, int ld, int m, const double* c,
                    const double* theta, int k, T* x, T* ax, double* norms2, double* ws,
                    hipStream_t s) {
    SI_CHECK(m >= 1 && m <= 64 && k >= 1 && k <= 8 && k <= m, kErrInvalidArgument,
             "rr_update: need k <= 8, k <= m <= 64");
    constexpr int R = sizeof(T) == 8 ? 32 : 64;
    int64_t blocks = ceil_div(n, (int64_t)R);
    if (blocks > kRrBlocks) blocks = kRrBlocks;
    if (m <= 16)
        hipLaunchKernelGGL((k_rr_update<T, 16>), dim3(blocks), dim3(256), 0, s, n, y, ly, ld, m, c, theta, k, x, ax, ws);
    else if (m <= 32)
        hipLaunchKernelGGL((k_rr_update<T, 32>), dim3(blocks), dim3(256), 0, s, n, y, ly, ld, m, c, theta, k, x, ax, ws);
    else
        hipLaunchKernelGGL((k_rr_update<T, 64>), dim3(blocks), dim3(256), 0, s, n, y, ly, ld, m, c, theta, k, x, ax, ws);
    SI_LAUNCH_CHECK();
    reduce_partials(ws, (int)blocks, 2 * k, norms2, s);
}

template <typename T>
void rr_update_seg_impl(const SegWork& tiles, const T* y, const T* ly, int ld, int m,
                        const double* c, const double* theta, int k, T* x, T* ax,
                        double* norms2, double* ws, hipStream_t s, const int32_t* gate) {
    SI_CHECK(m >= 1 && m <= 64 && k >= 1 && k <= 8 && k <= m, kErrInvalidArgument,
             "rr_update: need k <= 8, k <= m <= 64");
    const dim3 grid((unsigned)tiles.nblk);
    if (m <= 16)
        hipLaunchKernelGGL((k_rr_update_seg<T, 16>), grid, dim3(256), 0, s, tiles.blk_rows, tiles.blk_seg, y, ly, ld, m, c, theta, k, x, ax, ws, gate);
    else if (m <= 32)
        hipLaunchKernelGGL((k_rr_update_seg<T, 32>), grid, dim3(256), 0, s, tiles.blk_rows, tiles.blk_seg, y, ly, ld, m, c, theta, k, x, ax, ws, gate);
    else
        hipLaunchKernelGGL((k_rr_update_seg<T, 64>), grid, dim3(256), 0, s, tiles.blk_rows, tiles.blk_seg, y, ly, ld, m, c, theta, k, x, ax, ws, gate);
    SI_LAUNCH_CHECK();
    reduce_partials_seg(ws, tiles.seg_blk, tiles.nseg, 2 * k, norms2, s, gate);
}

// ------------------------------------------------------------------ elementwise
template <typename T>
__global__ void k_residual(int64_t n, const T* __restrict__ ax, const T* __restrict__ x, int ld,
                           int k, const double* __restrict__ theta, T* __restrict__ r, int ldr) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * k) return;
    const int64_t i = e / k;
    const int c = (int)(e % k);
    r[i * ldr + c] = ax[i * ld + c] - (T)theta[c] * x[i * ld + c];
}

// out = alpha * a + beta * b on the first w columns (common leading dimension)
template <typename T>
__global__ void k_axpby(int64_t n, int w, int ld, T alpha, const T* __restrict__ a, T beta,
                        const T* __restrict__ b, T* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * w) return;
    const int64_t o = (e / w) * ld + e % w;
    out[o] = alpha * a[o] + beta * b[o];
}

// Row i draws the normals of original node perm[i], so a reordered solve
// starts from the same block as an unordered one.  The draws are FP64; an FP32
// block holds them rounded.
template <typename T>
__global__ void k_fill_normal(T* __restrict__ x, int64_t n, int k, int ld, uint64_t seed,
                              uint32_t stream, const int32_t* __restrict__ perm) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * k) return;
    const int64_t i = e / k;
    const int c = (int)(e % k);
    const int64_t orig = perm ? (int64_t)perm[i] : i;
    x[i * ld + c] = (T)normal_at(seed, stream, (uint64_t)(orig * k + c));
}

template <typename T>
__global__ void k_gather_rows(const T* __restrict__ src, int lds, T* __restrict__ dst, int ldd,
                              int64_t n, int k, const int32_t* __restrict__ idx) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * k) return;
    const int64_t i = e / k;
    const int c = (int)(e % k);
    const int64_t r = idx ? (int64_t)idx[i] : i;
    dst[i * ldd + c] = src[r * lds + c];
}

// FP32 start blocks: thread per (row, column pair); row i draws the pairs of
// original node perm[i] (counter orig * ceil(k/2) + pair).
__global__ void k_fill_normal_pairs(float* __restrict__ x, int64_t n, int k, int ld, uint64_t seed,
                                    uint32_t stream, const int32_t* __restrict__ perm) {
    const int kp = (k + 1) / 2;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * kp) return;
    const int64_t i = e / kp;
    const int p = (int)(e % kp);
    const int64_t orig = perm ? (int64_t)perm[i] : i;
    float z0, z1;
    normal_pair_f32(seed, stream, (uint64_t)(orig * kp + p), z0, z1);
    float* row = x + i * ld + 2 * p;
    row[0] = z0;
    if (2 * p + 1 < k) row[1] = z1;
}

template <typename A, typename B>
__global__ void k_convert(const A* __restrict__ a, B* __restrict__ b, int64_t count) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) b[e] = (B)a[e];
}

template <typename A, typename B>
__global__ void k_copy_cols(const A* __restrict__ a, int lda, B* __restrict__ b, int ldb,
                            int64_t n, int k) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * k) return;
    const int64_t i = e / k;
    const int c = (int)(e % k);
    b[i * ldb + c] = (B)a[i * lda + c];
}

template <typename T>
__global__ void k_scale_cols(T* __restrict__ x, int ld, int64_t n, int k,
                             const double* __restrict__ scale) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * k) return;
    const int64_t i = e / k;
    const int c = (int)(e % k);
    x[i * ld + c] = (T)((double)x[i * ld + c] * scale[c]);
}

__global__ void k_average(double* __restrict__ x, const double* __restrict__ t, int64_t count) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) x[e] = (x[e] + t[e]) / 2.0;
}

template <typename T>
__global__ void k_set_column(T* __restrict__ x, int ld, int64_t n, int col,
                             const double* __restrict__ v, double scale) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i * ld + col] = (T)(v[i] * scale);
}

inline dim3 grid1(int64_t count) { return dim3((unsigned)ceil_div(count, 256)); }

}  // namespace

size_t gram_workspace_doubles(int64_t n, int p, int q) {
    int64_t nb = ceil_div(n, 64);
    if (nb > kGramBlocks) nb = kGramBlocks;
    if (nb < 1) nb = 1;
    return (size_t)nb * p * q;
}

void gram_f32(int64_t n, const float* a, int lda, int p, const float* b, int ldb, int q,
              double* g, double* ws, hipStream_t s) {
    gram_impl<float>(n, a, lda, p, b, ldb, q, nullptr, 0, q, g, ws, s);
}
void gram_f64(int64_t n, const double* a, int lda, int p, const double* b, int ldb, int q,
              double* g, double* ws, hipStream_t s) {
    gram_impl<double>(n, a, lda, p, b, ldb, q, nullptr, 0, q, g, ws, s);
}
void rr_update_f32(int64_t n, const float* y, const float* ly, int ld, int m, const double* c,
                   const double* theta, int k, float* x, float* ax, double* norms2, double* ws,
                   hipStream_t s) {
    rr_update_impl<float>(n, y, ly, ld, m, c, theta, k, x, ax, norms2, ws, s);
}
void rr_update_f64(int64_t n, const double* y, const double* ly, int ld, int m, const double* c,
                   const double* theta, int k, double* x, double* ax, double* norms2, double* ws,
                   hipStream_t s) {
    rr_update_impl<double>(n, y, ly, ld, m, c, theta, k, x, ax, norms2, ws, s);
}
size_t rr_update_workspace_doubles(int64_t n, int k) {
    int64_t nb = ceil_div(n, (int64_t)32);
    if (nb > kRrBlocks) nb = kRrBlocks;
    return (size_t)nb * 2 * k;
}

void gram2_seg_f32(const SegWork& blocks, const float* a, int ld, int p, const float* b2, int ld2,
                   double* g, double* ws, hipStream_t s, const int32_t* gate) {
    gram_impl<float>(0, a, ld, p, a, ld, p, b2, ld2, 2 * p, g, ws, s, &blocks, gate);
}
void gram2_seg_f64(const SegWork& blocks, const double* a, int ld, int p, const double* b2,
                   int ld2, double* g, double* ws, hipStream_t s, const int32_t* gate) {
    gram_impl<double>(0, a, ld, p, a, ld, p, b2, ld2, 2 * p, g, ws, s, &blocks, gate);
}
void rr_update_seg_f32(const SegWork& tiles, const float* y, const float* ly, int ld, int m,
                       const double* c, const double* theta, int k, float* x, float* ax,
                       double* norms2, double* ws, hipStream_t s, const int32_t* gate) {
    rr_update_seg_impl<float>(tiles, y, ly, ld, m, c, theta, k, x, ax, norms2, ws, s, gate);
}
void rr_update_seg_f64(const SegWork& tiles, const double* y, const double* ly, int ld, int m,
                       const double* c, const double* theta, int k, double* x, double* ax,
                       double* norms2, double* ws, hipStream_t s, const int32_t* gate) {
    rr_update_seg_impl<double>(tiles, y, ly, ld, m, c, theta, k, x, ax, norms2, ws, s, gate);
}
void gram2_f32(int64_t n, const float* a, int lda, int p, const float* b, int ldb, int q1,
               const float* b2, int ldb2, int q2, double* g, double* ws, hipStream_t s) {
    gram_impl<float>(n, a, lda, p, b, ldb, q1, b2, ldb2, q1 + q2, g, ws, s);
}
void gram2_f64(int64_t n, const double* a, int lda, int p, const double* b, int ldb, int q1,
               const double* b2, int ldb2, int q2, double* g, double* ws, hipStream_t s) {
    gram_impl<double>(n, a, lda, p, b, ldb, q1, b2, ldb2, q1 + q2, g, ws, s);
}

void block_mul_f32(int64_t n, const float* a, int lda, int p, const double* c, int q,
                   const float* b, int ldb, int p2, const double* d, float* out, int ldo,
                   hipStream_t s) {
    block_mul_impl<float>(n, a, lda, p, c, q, b, ldb, p2, d, out, ldo, s);
}
void block_mul_f64(int64_t n, const double* a, int lda, int p, const double* c, int q,
                   const double* b, int ldb, int p2, const double* d, double* out, int ldo,
                   hipStream_t s) {
    block_mul_impl<double>(n, a, lda, p, c, q, b, ldb, p2, d, out, ldo, s);
}

void residual_f32(int64_t n, const float* ax, const float* x, int ld, int k, const double* theta,
                  float* r, int ldr, hipStream_t s) {
    hipLaunchKernelGGL((k_residual<float>), grid1(n * k), dim3(256), 0, s, n, ax, x, ld, k, theta, r, ldr);
    SI_LAUNCH_CHECK();
}
void residual_f64(int64_t n, const double* ax, const double* x, int ld, int k,
                  const double* theta, double* r, int ldr, hipStream_t s) {
    hipLaunchKernelGGL((k_residual<double>), grid1(n * k), dim3(256), 0, s, n, ax, x, ld, k, theta, r, ldr);
    SI_LAUNCH_CHECK();
}

void gather_rows_f64(const double* src, int lds, double* dst, int ldd, int64_t n, int k,
                     const int32_t* idx, hipStream_t s) {
    hipLaunchKernelGGL((k_gather_rows<double>), grid1(n * k), dim3(256), 0, s, src, lds, dst, ldd, n, k, idx);
    SI_LAUNCH_CHECK();
}
void gather_rows_f32(const float* src, int lds, float* dst, int ldd, int64_t n, int k,
                     const int32_t* idx, hipStream_t s) {
    hipLaunchKernelGGL((k_gather_rows<float>), grid1(n * k), dim3(256), 0, s, src, lds, dst, ldd, n, k, idx);
    SI_LAUNCH_CHECK();
}

void fill_normal_f64(double* x, int64_t n, int k, int ld, uint64_t seed, uint32_t stream,
                     hipStream_t s, const int32_t* perm) {
    hipLaunchKernelGGL((k_fill_normal<double>), grid1(n * k), dim3(256), 0, s, x, n, k, ld, seed,
                       stream, perm);
    SI_LAUNCH_CHECK();
}
void fill_normal_rounded_f32(float* x, int64_t n, int k, int ld, uint64_t seed, uint32_t stream,
                             hipStream_t s, const int32_t* perm) {
    hipLaunchKernelGGL((k_fill_normal<float>), grid1(n * k), dim3(256), 0, s, x, n, k, ld, seed,
                       stream, perm);
    SI_LAUNCH_CHECK();
}
void fill_normal_f32(float* x, int64_t n, int k, int ld, uint64_t seed, uint32_t stream,
                     hipStream_t s, const int32_t* perm) {
    const int64_t pairs = n * ((k + 1) / 2);
    hipLaunchKernelGGL(k_fill_normal_pairs, grid1(pairs), dim3(256), 0, s, x, n, k, ld, seed,
                       stream, perm);
    SI_LAUNCH_CHECK();
}
void axpby_cols_f32(int64_t n, int w, int ld, float alpha, const float* a, float beta,
                    const float* b, float* out, hipStream_t s) {
    hipLaunchKernelGGL((k_axpby<float>), grid1(n * w), dim3(256), 0, s, n, w, ld, alpha, a, beta,
                       b, out);
    SI_LAUNCH_CHECK();
}
void axpby_cols_f64(int64_t n, int w, int ld, double alpha, const double* a, double beta,
                    const double* b, double* out, hipStream_t s) {
    hipLaunchKernelGGL((k_axpby<double>), grid1(n * w), dim3(256), 0, s, n, w, ld, alpha, a, beta,
                       b, out);
    SI_LAUNCH_CHECK();
}
void convert_f64_to_f32(const double* a, float* b, int64_t count, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<double, float>), grid1(count), dim3(256), 0, s, a, b, count);
    SI_LAUNCH_CHECK();
}
void convert_f32_to_f64(const float* a, double* b, int64_t count, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<float, double>), grid1(count), dim3(256), 0, s, a, b, count);
    SI_LAUNCH_CHECK();
}
void copy_cols_f64(const double* a, int lda, double* b, int ldb, int64_t n, int k, hipStream_t s) {
    hipLaunchKernelGGL((k_copy_cols<double, double>), grid1(n * k), dim3(256), 0, s, a, lda, b, ldb, n, k);
    SI_LAUNCH_CHECK();
}
void copy_cols_f64_to_f32(const double* a, int lda, float* b, int ldb, int64_t n, int k,
                          hipStream_t s) {
    hipLaunchKernelGGL((k_copy_cols<double, float>), grid1(n * k), dim3(256), 0, s, a, lda, b, ldb, n, k);
    SI_LAUNCH_CHECK();
}
void copy_cols_f32(const float* a, int lda, float* b, int ldb, int64_t n, int k, hipStream_t s) {
    hipLaunchKernelGGL((k_copy_cols<float, float>), grid1(n * k), dim3(256), 0, s, a, lda, b, ldb, n, k);
    SI_LAUNCH_CHECK();
}
void scale_cols_f32(float* x, int ld, int64_t n, int k, const double* scale, hipStream_t s) {
    hipLaunchKernelGGL((k_scale_cols<float>), grid1(n * k), dim3(256), 0, s, x, ld, n, k, scale);
    SI_LAUNCH_CHECK();
}
void scale_cols_f64(double* x, int ld, int64_t n, int k, const double* scale, hipStream_t s) {
    hipLaunchKernelGGL((k_scale_cols<double>), grid1(n * k), dim3(256), 0, s, x, ld, n, k, scale);
    SI_LAUNCH_CHECK();
}
void average_f64(double* x, const double* t, int64_t count, hipStream_t s) {
    hipLaunchKernelGGL(k_average, grid1(count), dim3(256), 0, s, x, t, count);
    SI_LAUNCH_CHECK();
}
void set_column_f64(double* x, int ld, int64_t n, int col, const double* v, double scale,
                    hipStream_t s) {
    hipLaunchKernelGGL((k_set_column<double>), grid1(n), dim3(256), 0, s, x, ld, n, col, v, scale);
    SI_LAUNCH_CHECK();
}
void set_column_f32(float* x, int ld, int64_t n, int col, const double* v, double scale,
                    hipStream_t s) {
    hipLaunchKernelGGL((k_set_column<float>), grid1(n), dim3(256), 0, s, x, ld, n, col, v, scale);
    SI_LAUNCH_CHECK();
}

}  // namespace si
