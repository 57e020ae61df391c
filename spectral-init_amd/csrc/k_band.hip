// k_band.hip -- banded Cholesky factorisation of M = L + shift I and solves
// with it, in the device's node order: the operator of the shift-and-invert
// LOBPCG (src/solvers/sinv.rs:106-189), where the reference factorises M
// with a sparse Cholesky (sinv.rs:51-73, 125) and applies M^-1 per iteration
// (sinv.rs:139-153).
//
// Storage: the lower band of bandwidth b, row-major, w = b + 1 doubles a row;
// entry (i, j), i - b <= j <= i, lives at B[i * w + (j - i + b)].  Graphs of
// 4096 nodes or more are stored in reverse Cuthill-McKee order, which keeps b
// small for path-, ring- and grid-like graphs; below 4096 nodes the band may
// be the whole lower triangle (a dense factorisation).  A Cholesky factor of a
// band matrix stays inside the band, so the factor overwrites M in place.
//
// Factorisation, per panel of 64 columns [j0, j1): k_band_diag factors the
// diagonal block in LDS (one workgroup), k_band_trsm solves the rows below it
// that the band reaches (one thread a row), k_band_syrk subtracts the panel's
// contribution from the trailing band (64 x 64 tiles staged through LDS).
// Solves: one workgroup per right-hand side sweeps the panels; the part of a
// row outside its panel is a wave-reduced dot product, the triangle inside
// the panel is solved by one wavefront with lane r holding row r.
//
// Dense fallback: when the band exceeds kBandMaxEntries / kBandMaxFlops (graphs
// with no narrow ordering, e.g. k-NN graphs) and n <= kSinvDenseMax, M is
// assembled as a full n x n matrix and factorised by rocSOLVER DPOTRF (a plain
// library Cholesky on the FP64 matrix cores).  A row-major n x k block Y is the
// column-major k x n matrix Y^T, so Y = M^-1 X is Y^T = X^T L^-T L^-1: two
// right-side rocBLAS DTRSM on the block in place, no transposes.
#include <rocsolver/rocsolver.h>

#include "si_host.hpp"
#include "si_rocblas.hpp"

namespace si {

namespace {

constexpr int kP = 64;  // panel width
bool g_force_dense = false;

// lower triangle of M = L + shift I, column-major n x n (zeroed beforehand)
__global__ void k_dense_fill(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                             const double* __restrict__ l64, int n, double shift,
                             double* __restrict__ A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    A[(int64_t)i * n + i] = 1.0 + shift;
    for (int32_t p = rp[i]; p < rp[i + 1]; ++p) {
        const int j = col[p];
        if (j < i) A[(int64_t)j * n + i] = l64[p];
    }
}

// y = scale * x for an n x k row-major block
__global__ void k_dense_copy(const double* __restrict__ x, int ldx, double* __restrict__ y,
                             int ldy, int64_t n, int k, double scale) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * k) return;
    const int64_t i = e / k;
    const int c = (int)(e % k);
    y[i * ldy + c] = scale * x[i * ldx + c];
}

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// largest i - j over the stored entries (i, j) with j < i
__global__ __launch_bounds__(256) void k_band_width(const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ col, int n,
                                                    int* __restrict__ out) {
    __shared__ int red[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int bw = 0;
    if (i < n)
        for (int32_t p = rp[i]; p < rp[i + 1]; ++p) bw = max(bw, i - col[p]);
    red[threadIdx.x] = bw;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, red[0]);
}

// band of M = I + shift I + off-diagonal entries of L (the band was zeroed)
__global__ void k_band_fill(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                            const double* __restrict__ l64, int n, int b, double shift,
                            double* __restrict__ B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t w = b + 1;
    B[i * w + b] = 1.0 + shift;
    for (int32_t p = rp[i]; p < rp[i + 1]; ++p) {
        const int j = col[p];
        if (j < i) B[i * w + (j - i + b)] = l64[p];
    }
}

// Cholesky of the diagonal block [j0, j0 + pn) in LDS (one workgroup).
// status[0] is set when a pivot is not positive (M not positive definite).
__global__ __launch_bounds__(256) void k_band_diag(double* __restrict__ B, int b, int j0, int pn,
                                                   int* __restrict__ status) {
    __shared__ double P[kP][kP + 1];
    const int t = threadIdx.x;
    const int64_t w = b + 1;
    for (int e = t; e < pn * pn; e += 256) {
        const int r = e / pn, q = e % pn;
        P[r][q] = (q <= r && q >= r - b) ? B[(int64_t)(j0 + r) * w + (q - r + b)] : 0.0;
    }
    for (int c = 0; c < pn; ++c) {
        __syncthreads();
        const double d = P[c][c];
        const bool ok = d > 0.0;  // false for NaN too
        const double sd = ok ? sqrt(d) : 1.0;
        if (!ok && t == 0) status[0] = 1;
        __syncthreads();
        for (int r = c + 1 + t; r < pn; r += 256) P[r][c] /= sd;
        if (t == 0) P[c][c] = sd;
        __syncthreads();
        const int m = pn - c - 1;
        for (int e = t; e < m * m; e += 256) {
            const int r = c + 1 + e / m, q = c + 1 + e % m;
            if (q <= r) P[r][q] -= P[r][c] * P[q][c];
        }
    }
    __syncthreads();
    for (int e = t; e < pn * pn; e += 256) {
        const int r = e / pn, q = e % pn;
        if (q <= r && q >= r - b) B[(int64_t)(j0 + r) * w + (q - r + b)] = P[r][q];
    }
}

// Rows i in [j1, j1 + nr) below the panel: x P^T = a, with a = M[i, j0 .. j1)
// (zero outside the band) and P the panel's factored diagonal block.
constexpr int kTrsmThreads = 128;
__global__ __launch_bounds__(kTrsmThreads) void k_band_trsm(double* __restrict__ B, int b, int j0,
                                                            int pn, int j1, int nr) {
    __shared__ double P[kP][kP + 1];
    __shared__ double xs[kP][kTrsmThreads];
    const int t = threadIdx.x;
    const int64_t w = b + 1;
    for (int e = t; e < pn * pn; e += kTrsmThreads) {
        const int r = e / pn, q = e % pn;
        P[r][q] = (q <= r && q >= r - b) ? B[(int64_t)(j0 + r) * w + (q - r + b)] : 0.0;
    }
    __syncthreads();
    const int i = j1 + blockIdx.x * kTrsmThreads + t;
    if (i >= j1 + nr) return;
    double* row = B + (int64_t)i * w;
    for (int q = 0; q < pn; ++q) {
        const int j = j0 + q;
        const bool in_band = j >= i - b;
        double a = in_band ? row[j - i + b] : 0.0;
        for (int p = 0; p < q; ++p) a -= xs[p][t] * P[q][p];
        a /= P[q][q];
        xs[q][t] = a;
        if (in_band) row[j - i + b] = a;
    }
}

// Trailing update M[i][q] -= sum_c L[i][j0+c] L[q][j0+c] for j1 <= q <= i <
// j1 + nr inside the band; workgroup (ti, tq) owns a 64 x 64 tile.
__global__ __launch_bounds__(256) void k_band_syrk(double* __restrict__ B, int b, int j0, int pn,
                                                   int j1, int nr) {
    __shared__ double LI[kP][kP + 1];
    __shared__ double LQ[kP][kP + 1];
    const int ti = blockIdx.x, tq = blockIdx.y;
    if (tq > ti) return;
    const int i0 = j1 + ti * kP, q0 = j1 + tq * kP;
    if (q0 + kP - 1 < i0 - b) return;  // the tile lies outside the band
    const int t = threadIdx.x;
    const int64_t w = b + 1;
    const int iend = j1 + nr;
    for (int e = t; e < kP * pn; e += 256) {
        const int r = e / pn, c = e % pn;
        const int j = j0 + c;
        const int ii = i0 + r, qq = q0 + r;
        LI[r][c] = (ii < iend && j >= ii - b) ? B[(int64_t)ii * w + (j - ii + b)] : 0.0;
        LQ[r][c] = (qq < iend && j >= qq - b) ? B[(int64_t)qq * w + (j - qq + b)] : 0.0;
    }
    __syncthreads();
    const int r = t & 63;
    const int qb = (t >> 6) * 16;
    double acc[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) acc[u] = 0.0;
    for (int c = 0; c < pn; ++c) {
        const double li = LI[r][c];
#pragma unroll
        for (int u = 0; u < 16; ++u) acc[u] += li * LQ[qb + u][c];
    }
    const int i = i0 + r;
    if (i >= iend) return;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int q = q0 + qb + u;
        if (q <= i && q >= i - b && q < iend) B[(int64_t)i * w + (q - i + b)] -= acc[u];
    }
}

// y = L^-1 x for column blockIdx.x (x, y: [n x k] with leading dims ldx, ldy)
__global__ __launch_bounds__(256) void k_band_fwd(const double* __restrict__ B, int n, int b,
                                                  const double* __restrict__ x, int ldx,
                                                  double* __restrict__ y, int ldy) {
    __shared__ double part[kP];
    const int c = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int64_t w = b + 1;
    for (int j0 = 0; j0 < n; j0 += kP) {
        const int pn = min(kP, n - j0);
        for (int r = wv; r < pn; r += 4) {
            const int i = j0 + r;
            const int lo = max(0, i - b);
            const double* row = B + (int64_t)i * w + (b - i);  // row[j] = L[i][j]
            double s = 0.0;
            for (int j = lo + lane; j < j0; j += 64) s += row[j] * y[(int64_t)j * ldy + c];
            s = wave_sum(s);
            if (lane == 0) part[r] = x[(int64_t)i * ldx + c] - s;
        }
        __syncthreads();
        if (wv == 0) {
            const int i = j0 + lane;
            const bool act = lane < pn;
            const double* row = B + (int64_t)(act ? i : j0) * w + (b - (act ? i : j0));
            double tr = act ? part[lane] : 0.0;
            for (int q = 0; q < pn; ++q) {
                if (lane == q) tr /= row[j0 + q];
                const double yq = __shfl(tr, q, 64);
                if (act && lane > q && j0 + q >= i - b) tr -= row[j0 + q] * yq;
            }
            if (act) y[(int64_t)i * ldy + c] = tr;
        }
        __syncthreads();
    }
}

// y = scale * L^-T y in place for column blockIdx.x (scale is +1 or -1, so
// scaling a stored value twice returns it exactly)
__global__ __launch_bounds__(256) void k_band_bwd(const double* __restrict__ B, int n, int b,
                                                  double* __restrict__ y, int ldy, double scale) {
    __shared__ double red[4][kP];
    const int c = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int64_t w = b + 1;
    const int last = ((n - 1) / kP) * kP;
    for (int j0 = last; j0 >= 0; j0 -= kP) {
        const int pn = min(kP, n - j0);
        const int j1 = j0 + pn;
        const int jend = min(n, j1 + b);  // rows that reach into the panel
        const int i = j0 + lane;
        double s = 0.0;
        for (int j = j1 + wv; j < jend; j += 4) {
            const double zj = y[(int64_t)j * ldy + c] * scale;
            if (lane < pn && i >= j - b) s += B[(int64_t)j * w + (i - j + b)] * zj;
        }
        red[wv][lane] = s;
        __syncthreads();
        if (wv == 0) {
            const bool act = lane < pn;
            double tr = act ? y[(int64_t)i * ldy + c] - (red[0][lane] + red[1][lane] +
                                                         red[2][lane] + red[3][lane])
                            : 0.0;
            for (int q = pn - 1; q >= 0; --q) {
                const double* rowq = B + (int64_t)(j0 + q) * w + (b - (j0 + q));
                if (lane == q) tr /= rowq[j0 + q];
                const double zq = __shfl(tr, q, 64);
                if (lane < q && i >= j0 + q - b) tr -= rowq[i] * zq;
            }
            if (act) y[(int64_t)i * ldy + c] = tr * scale;
        }
        __syncthreads();
    }
}

}  // namespace

int band_width(const DevLaplacian& L, hipStream_t s) {
    const int n = (int)L.n;
    DevBuf<int> bw(1);
    SI_HIP(hipMemsetAsync(bw.get(), 0, sizeof(int), s));
    hipLaunchKernelGGL(k_band_width, dim3(ceil_div(n, 256)), dim3(256), 0, s, L.rowptr, L.col, n,
                       bw.get());
    SI_LAUNCH_CHECK();
    int h = 0;
    to_host(&h, bw.get(), 1, s);
    return h;
}

namespace {

bool dense_factor(const DevLaplacian& L, double shift, BandFactor& F, hipStream_t s) {
    const int64_t n = L.n;
    if (n < 1 || n > kSinvDenseMax) return false;
    const size_t bytes = sizeof(double) * (size_t)n * (size_t)n;
    size_t free_b = 0, total_b = 0;
    SI_HIP(hipMemGetInfo(&free_b, &total_b));
    if ((double)bytes > kSinvDenseMemFrac * (double)free_b) return false;
    F.n = (int)n;
    F.b = (int)n - 1;
    F.dense = true;
    try {
        F.B.ensure((size_t)n * (size_t)n);
    } catch (const Error&) {  // out of device memory: no factor, as a failed Cholesky
        (void)hipGetLastError();
        F.dense = false;
        return false;
    }
    SI_HIP(hipMemsetAsync(F.B.get(), 0, bytes, s));
    hipLaunchKernelGGL(k_dense_fill, dim3(ceil_div(n, (int64_t)256)), dim3(256), 0, s, L.rowptr,
                       L.col, L.l64, (int)n, shift, F.B.get());
    SI_LAUNCH_CHECK();
    DevBuf<rocblas_int> info(1);
    rb_check(rocsolver_dpotrf(rb_handle(s), rocblas_fill_lower, (rocblas_int)n, F.B.get(),
                              (rocblas_int)n, info.get()),
             "rocsolver_dpotrf");
    rocblas_int h = 0;
    to_host(&h, info.get(), 1, s);
    return h == 0;
}

}  // namespace

void sinv_force_dense(bool on) { g_force_dense = on; }

bool band_factor(const DevLaplacian& L, double shift, BandFactor& F, hipStream_t s) {
    const int64_t n64 = L.n;
    if (n64 < 1 || n64 > INT32_MAX) return false;
    const int n = (int)n64;
    F.dense = false;
    if (g_force_dense) return dense_factor(L, shift, F, s);
    const int b = band_width(L, s);
    const double entries = (double)n * (double)(b + 1);
    const double flops = (double)n * (double)b * (double)b;
    if (entries > kBandMaxEntries || flops > kBandMaxFlops) return dense_factor(L, shift, F, s);
    F.n = n;
    F.b = b;
    F.B.ensure((size_t)n * (size_t)(b + 1));
    DevBuf<int> status(1);
    SI_HIP(hipMemsetAsync(F.B.get(), 0, sizeof(double) * (size_t)n * (size_t)(b + 1), s));
    SI_HIP(hipMemsetAsync(status.get(), 0, sizeof(int), s));
    hipLaunchKernelGGL(k_band_fill, dim3(ceil_div(n, 256)), dim3(256), 0, s, L.rowptr, L.col,
                       L.l64, n, b, shift, F.B.get());
    SI_LAUNCH_CHECK();
    for (int j0 = 0; j0 < n; j0 += kP) {
        const int pn = std::min(kP, n - j0);
        const int j1 = j0 + pn;
        const int nr = std::min(n, j1 + b) - j1;  // rows below that the band reaches
        hipLaunchKernelGGL(k_band_diag, dim3(1), dim3(256), 0, s, F.B.get(), b, j0, pn,
                           status.get());
        if (nr > 0) {
            hipLaunchKernelGGL(k_band_trsm, dim3(ceil_div(nr, kTrsmThreads)), dim3(kTrsmThreads),
                               0, s, F.B.get(), b, j0, pn, j1, nr);
            const int tiles = ceil_div(nr, kP);
            hipLaunchKernelGGL(k_band_syrk, dim3(tiles, tiles), dim3(256), 0, s, F.B.get(), b, j0,
                               pn, j1, nr);
        }
    }
    SI_LAUNCH_CHECK();
    int h = 0;
    to_host(&h, status.get(), 1, s);
    return h == 0;
}

void band_solve(const BandFactor& F, const double* x, int ldx, double* y, int ldy, int k,
                double scale, hipStream_t s) {
    SI_CHECK(scale == 1.0 || scale == -1.0, kErrInvalidArgument, "band_solve: scale must be +-1");
    SI_CHECK(k >= 1, kErrInvalidArgument, "band_solve: k < 1");
    if (F.dense) {
        SI_CHECK(ldy >= k && x != y, kErrInvalidArgument, "band_solve: bad dense solve operands");
        const int64_t nk = (int64_t)F.n * k;
        hipLaunchKernelGGL(k_dense_copy, dim3(ceil_div(nk, (int64_t)256)), dim3(256), 0, s, x,
                           ldx, y, ldy, (int64_t)F.n, k, scale);
        SI_LAUNCH_CHECK();
        const double one = 1.0;
        rocblas_handle h = rb_handle(s);
        // Y^T <- Y^T L^-T, then Y^T <- Y^T L^-1 (Y^T: column-major k x n, ld ldy)
        rb_check(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                               rocblas_diagonal_non_unit, k, F.n, &one, F.B.get(), F.n, y, ldy),
                 "rocblas_dtrsm (L^T)");
        rb_check(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_none,
                               rocblas_diagonal_non_unit, k, F.n, &one, F.B.get(), F.n, y, ldy),
                 "rocblas_dtrsm (L)");
        return;
    }
    hipLaunchKernelGGL(k_band_fwd, dim3(k), dim3(256), 0, s, (const double*)F.B.get(), F.n, F.b,
                       x, ldx, y, ldy);
    hipLaunchKernelGGL(k_band_bwd, dim3(k), dim3(256), 0, s, (const double*)F.B.get(), F.n, F.b,
                       y, ldy, scale);
    SI_LAUNCH_CHECK();
}

}  // namespace si
