// k_dense.hip -- dense symmetric eigensolver for the k smallest eigenpairs of a
// Laplacian of up to kDenseMax nodes: the device form of dense_evd
// (src/solvers/dense.rs:9-46), used by escalation levels 0 and 5
// (src/solvers/mod.rs:83-109, 211-223) and by si_dense_evd.
//
//   1. densify L (FP64, row-major n x n, unit diagonal);
//   2. Householder tridiagonalisation in n - 2 steps.  Step j: k_house forms
//      the reflector of row j's tail (stored over that tail), k_symv computes
//      p = 2 B v for the trailing block B, k_rank2 applies
//      B -= v w^T + w v^T with w = p - (v.p) v;
//   3. multisection on Sturm counts for the k smallest eigenvalues of the
//      tridiagonal matrix (one workgroup per eigenvalue, 256 points a round);
//   4. inverse iteration for their tridiagonal eigenvectors (LU of T - lambda I
//      with partial pivoting, three solves, Gram-Schmidt inside clusters of
//      close eigenvalues, the scheme of LAPACK's dstein);
//   5. back-transformation x = H_0 H_1 ... H_{n-3} y (one workgroup per vector).
// Every reduction has a fixed order, so results are bit-reproducible.
#include "si_host.hpp"

namespace si {

namespace {

constexpr int kT = 256;
constexpr int kWaves = kT / 64;
constexpr int kRank2Rows = 8;
constexpr double kEps = 2.220446049250313e-16;

// Block-wide sum: wave butterfly, then the kWaves wave totals in order.  Every
// thread returns the same value.
__device__ double block_sum(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();  // earlier readers of red are done
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < kWaves; ++w) r += red[w];
    return r;
}

// one thread per row: A[i][:] = I[i][:] + off-diagonal entries of L
__global__ void k_densify(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                          const double* __restrict__ l64, int n, double* __restrict__ a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double* ai = a + (int64_t)i * n;
    for (int j = 0; j < n; ++j) ai[j] = (i == j) ? 1.0 : 0.0;
    for (int32_t p = rp[i]; p < rp[i + 1]; ++p)
        if (col[p] != i) ai[col[p]] += l64[p];
}

// Reflector of step j (one workgroup).  x = A[j][j+1..n-1], alpha = -sign(x0)|x|,
// v = (x - alpha e1) / |x - alpha e1| written over x; d[j] = A[j][j],
// e[j] = alpha.  vn[j] = 1 when a reflection is applied, 0 when x is already
// a multiple of e1 (then e[j] = x0).
__global__ __launch_bounds__(kT) void k_house(double* __restrict__ a, int n, int j,
                                              double* __restrict__ d, double* __restrict__ e,
                                              double* __restrict__ vn) {
    __shared__ double red[kWaves];
    const int t = threadIdx.x;
    double* x = a + (int64_t)j * n + (j + 1);
    const int m = n - j - 1;
    const double x0 = x[0];  // every thread reads x0 before block_sum's barriers
    double s = 0.0;
    for (int q = 1 + t; q < m; q += kT) s += x[q] * x[q];
    const double tail2 = block_sum(s, red);
    if (t == 0) d[j] = a[(int64_t)j * n + j];
    if (tail2 == 0.0) {
        if (t == 0) { e[j] = x0; vn[j] = 0.0; }
        return;
    }
    const double nx = sqrt(x0 * x0 + tail2);
    const double alpha = x0 >= 0.0 ? -nx : nx;
    const double v0 = x0 - alpha;
    const double inv = 1.0 / sqrt(v0 * v0 + tail2);
    for (int q = t; q < m; q += kT) x[q] = (q == 0 ? v0 : x[q]) * inv;
    if (t == 0) { e[j] = alpha; vn[j] = 1.0; }
}

// p = 2 B v over the trailing block B = A[j+1.., j+1..]; one wavefront per row
__global__ __launch_bounds__(kT) void k_symv(const double* __restrict__ a, int n, int j,
                                             const double* __restrict__ vn,
                                             double* __restrict__ p) {
    if (vn[j] == 0.0) return;
    const int m = n - j - 1;
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (r >= m) return;
    const double* brow = a + (int64_t)(j + 1 + r) * n + (j + 1);
    const double* v = a + (int64_t)j * n + (j + 1);
    double s = 0.0;
    for (int q = lane; q < m; q += 64) s += brow[q] * v[q];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) p[r] = 2.0 * s;
}

// B -= v w^T + w v^T, w = p - (v.p) v.  Every workgroup computes v.p in the
// same order, then updates its kRank2Rows rows of B.
__global__ __launch_bounds__(kT) void k_rank2(double* __restrict__ a, int n, int j,
                                              const double* __restrict__ vn,
                                              const double* __restrict__ p) {
    __shared__ double red[kWaves];
    if (vn[j] == 0.0) return;
    const int m = n - j - 1;
    const double* v = a + (int64_t)j * n + (j + 1);
    double s = 0.0;
    for (int q = threadIdx.x; q < m; q += kT) s += v[q] * p[q];
    const double kk = block_sum(s, red);
    for (int rr = 0; rr < kRank2Rows; ++rr) {
        const int r = blockIdx.x * kRank2Rows + rr;
        if (r >= m) break;
        const double vr = v[r];
        const double wr = p[r] - kk * vr;
        double* brow = a + (int64_t)(j + 1 + r) * n + (j + 1);
        for (int q = threadIdx.x; q < m; q += kT) {
            const double vq = v[q];
            const double wq = p[q] - kk * vq;
            brow[q] -= vr * wq + wr * vq;
        }
    }
}

// the last 2 x 2 block of the tridiagonal matrix
__global__ void k_tail(const double* __restrict__ a, int n, double* __restrict__ d,
                       double* __restrict__ e) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (n >= 2) {
        d[n - 2] = a[(int64_t)(n - 2) * n + (n - 2)];
        e[n - 2] = a[(int64_t)(n - 2) * n + (n - 1)];
    }
    d[n - 1] = a[(int64_t)(n - 1) * n + (n - 1)];
}

// e2[i] = e[i]^2 and the scale of T (max Gershgorin radius bound), one workgroup
__global__ __launch_bounds__(kT) void k_tri_prep(const double* __restrict__ d,
                                                 const double* __restrict__ e, int n,
                                                 double* __restrict__ e2,
                                                 double* __restrict__ bounds) {
    __shared__ double s_lo[kT], s_hi[kT];
    const int t = threadIdx.x;
    double lo = INFINITY, hi = -INFINITY;
    for (int i = t; i < n; i += kT) {
        if (i < n - 1) e2[i] = e[i] * e[i];
        const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
        lo = fmin(lo, d[i] - r);
        hi = fmax(hi, d[i] + r);
    }
    s_lo[t] = lo;
    s_hi[t] = hi;
    __syncthreads();
    for (int st = kT / 2; st > 0; st >>= 1) {
        if (t < st) {
            s_lo[t] = fmin(s_lo[t], s_lo[t + st]);
            s_hi[t] = fmax(s_hi[t], s_hi[t + st]);
        }
        __syncthreads();
    }
    if (t == 0) {
        bounds[0] = s_lo[0];
        bounds[1] = s_hi[0];
        bounds[2] = fmax(fmax(fabs(s_lo[0]), fabs(s_hi[0])), 1e-300);  // ||T|| bound
    }
}

// Number of eigenvalues of the tridiagonal (d, e) below x: the negative pivots
// of the LDL^T factorisation of T - x I (e2[i] = e[i]^2; a pivot in
// (-tiny, 0] is replaced by -tiny).
__device__ int sturm_count(const double* d, const double* e2, int n, double x, double tiny) {
    int c = 0;
    double q = d[0] - x;
    for (int i = 0;; ++i) {
        if (q <= 0.0) {
            ++c;
            if (q > -tiny) q = -tiny;
        }
        if (i + 1 >= n) break;
        q = (d[i + 1] - x) - e2[i] / q;
    }
    return c;
}

// Workgroup j: the j-th smallest eigenvalue.  Each round evaluates the Sturm
// count at kT evenly spaced interior points of [lo, hi] and keeps the
// sub-interval where the count first exceeds j; rounds stop once
// hi - lo <= 2 eps (|lo| + |hi|) + 2 eps ||T||.
__global__ __launch_bounds__(kT) void k_bisect(const double* __restrict__ d,
                                               const double* __restrict__ e2, int n,
                                               const double* __restrict__ bounds,
                                               double* __restrict__ w) {
    __shared__ double s_lo, s_hi;
    __shared__ int s_first;
    const int j = blockIdx.x;
    const int t = threadIdx.x;
    const double anorm = bounds[2];
    const double tiny = kEps * kEps * anorm + 1e-300;
    const double abstol = 2.0 * kEps * anorm;
    if (t == 0) {
        s_lo = bounds[0] - abstol;
        s_hi = bounds[1] + abstol;
    }
    __syncthreads();
    for (int round = 0; round < 64; ++round) {
        const double lo = s_lo, hi = s_hi;
        if (hi - lo <= 2.0 * kEps * (fabs(lo) + fabs(hi)) + abstol) break;
        const double step = (hi - lo) / (double)(kT + 1);
        const double x = lo + step * (double)(t + 1);
        const bool above = sturm_count(d, e2, n, x, tiny) > j;
        if (t == 0) s_first = kT;
        __syncthreads();
        if (above) atomicMin(&s_first, t);
        __syncthreads();
        if (t == 0) {
            const int f = s_first;
            if (f == kT) {
                s_lo = lo + step * (double)kT;
            } else {
                s_hi = lo + step * (double)(f + 1);
                if (f > 0) s_lo = lo + step * (double)f;
            }
        }
        __syncthreads();
    }
    if (t == 0) w[j] = 0.5 * (s_lo + s_hi);
}

// Inverse iteration for the k tridiagonal eigenvectors (one workgroup).  The
// LU factorisation and the solves run on thread 0 (O(n) each); Gram-Schmidt
// against earlier vectors of the same cluster (|w_q - w_j| <= 1e-3 ||T||) and
// the normalisation are block-parallel.  y is n x k column-major.
__global__ __launch_bounds__(kT) void k_invit(const double* __restrict__ d,
                                              const double* __restrict__ e, int n, int k,
                                              const double* __restrict__ w,
                                              const double* __restrict__ bounds,
                                              double* __restrict__ lu, int* __restrict__ piv,
                                              double* __restrict__ y, uint64_t seed) {
    __shared__ double red[kWaves];
    const int t = threadIdx.x;
    const double anorm = bounds[2];
    const double pivmin = kEps * anorm;  // replaces zero pivots
    double* dl = lu;          // multipliers
    double* u0 = lu + n;      // U diagonal
    double* u1 = lu + 2 * n;  // U first super-diagonal
    double* u2 = lu + 3 * n;  // U second super-diagonal (non-zero after a swap)
    for (int j = 0; j < k; ++j) {
        double* yj = y + (int64_t)j * n;
        for (int i = t; i < n; i += kT) yj[i] = normal_at(seed, kStreamDenseInvit, (uint64_t)j * n + i);
        if (t == 0) {
            const double lam = w[j];
            double cp = d[0] - lam;          // pivot candidate of the current row
            double cq = n > 1 ? e[0] : 0.0;  // its super-diagonal
            for (int i = 0; i + 1 < n; ++i) {
                const double sub = e[i];
                const double dn = d[i + 1] - lam;
                const double en = i + 2 < n ? e[i + 1] : 0.0;
                if (fabs(cp) >= fabs(sub)) {
                    const double pv = fabs(cp) < pivmin ? (cp < 0.0 ? -pivmin : pivmin) : cp;
                    const double l = sub / pv;
                    u0[i] = pv; u1[i] = cq; u2[i] = 0.0; dl[i] = l; piv[i] = 0;
                    cp = dn - l * cq;
                    cq = en;
                } else {
                    const double l = cp / sub;
                    u0[i] = sub; u1[i] = dn; u2[i] = en; dl[i] = l; piv[i] = 1;
                    cp = cq - l * dn;
                    cq = -l * en;
                }
            }
            u0[n - 1] = fabs(cp) < pivmin ? (cp < 0.0 ? -pivmin : pivmin) : cp;
        }
        __syncthreads();
        for (int sweep = 0; sweep < 3; ++sweep) {
            if (t == 0) {
                for (int i = 0; i + 1 < n; ++i) {
                    if (piv[i]) { const double tmp = yj[i]; yj[i] = yj[i + 1]; yj[i + 1] = tmp; }
                    yj[i + 1] -= dl[i] * yj[i];
                }
                for (int i = n - 1; i >= 0; --i) {
                    double s = yj[i];
                    if (i + 1 < n) s -= u1[i] * yj[i + 1];
                    if (i + 2 < n) s -= u2[i] * yj[i + 2];
                    yj[i] = s / u0[i];
                }
                // scale down to keep later solves finite
                double mx = 0.0;
                for (int i = 0; i < n; ++i) mx = fmax(mx, fabs(yj[i]));
                const double sc = mx > 0.0 ? 1.0 / mx : 1.0;
                for (int i = 0; i < n; ++i) yj[i] *= sc;
            }
            __syncthreads();
            for (int q = 0; q < j; ++q) {
                if (fabs(w[q] - w[j]) > 1e-3 * anorm) continue;
                const double* yq = y + (int64_t)q * n;
                double s = 0.0;
                for (int i = t; i < n; i += kT) s += yq[i] * yj[i];
                const double c = block_sum(s, red);
                for (int i = t; i < n; i += kT) yj[i] -= c * yq[i];
            }
            double s = 0.0;
            for (int i = t; i < n; i += kT) s += yj[i] * yj[i];
            const double nrm = sqrt(block_sum(s, red));
            const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
            for (int i = t; i < n; i += kT) yj[i] *= inv;
            __syncthreads();
        }
    }
}

// Workgroup c: column c of y through H_{n-3}, ..., H_0, then written to
// out[i * ldo + c].  Thread t owns the rows i = t (mod kT) throughout, so the
// update of one reflector and the dot of the next need no extra barrier.
__global__ __launch_bounds__(kT) void k_backtransform(const double* __restrict__ a, int n,
                                                      const double* __restrict__ vn,
                                                      double* __restrict__ y,
                                                      double* __restrict__ out, int ldo) {
    __shared__ double red[kWaves];
    const int t = threadIdx.x;
    const int c = blockIdx.x;
    double* yc = y + (int64_t)c * n;
    for (int j = n - 3; j >= 0; --j) {
        if (vn[j] == 0.0) continue;
        const double* arow = a + (int64_t)j * n;  // v_j = arow[j+1 .. n-1]
        const int first = j + 1 + ((t - (j + 1)) % kT + kT) % kT;
        double s = 0.0;
        for (int i = first; i < n; i += kT) s += arow[i] * yc[i];
        const double dot2 = 2.0 * block_sum(s, red);
        for (int i = first; i < n; i += kT) yc[i] -= dot2 * arow[i];
    }
    for (int i = t; i < n; i += kT) out[(int64_t)i * ldo + c] = yc[i];
}

}  // namespace

// dense_evd (src/solvers/dense.rs:9-46) for kSmallMax < n <= kDenseMax:
// eigenvalues to the host, eigenvectors to evecs_dev (n x k row-major, the
// device's permuted row order).
void dense_evd_large(const DevLaplacian& L, int k, double* evals_host, double* evecs_dev,
                     hipStream_t s) {
    const int n = (int)L.n;
    SI_CHECK(n >= 1 && n <= kDenseMax, kErrUnsupported,
             "dense_evd: the device dense solver supports n <= 4096");
    SI_CHECK(k >= 1 && k <= n, kErrTooFewNodes, "dense_evd: k > n");
    SI_CHECK(k <= kSmallMax, kErrUnsupported, "dense_evd: k above 64 is not supported");
    double* a = ws_get<double>(kWsDenseA, (size_t)n * n);
    // d, e, e2, vn, p [5n] | lu [4n] | y [n k] | w [k] | bounds [4]
    double* aux = ws_get<double>(kWsDenseAux, (size_t)n * (9 + k) + k + 4);
    int* piv = ws_get<int>(kWsDensePiv, (size_t)n);
    double* d = aux;
    double* e = d + n;
    double* e2 = e + n;
    double* vn = e2 + n;
    double* p = vn + n;
    double* lu = p + n;
    double* y = lu + 4 * (size_t)n;
    double* w = y + (size_t)n * k;
    double* bounds = w + k;
    hipLaunchKernelGGL(k_densify, dim3(ceil_div(n, 256)), dim3(256), 0, s, L.rowptr, L.col, L.l64,
                       n, a);
    SI_LAUNCH_CHECK();
    for (int j = 0; j + 2 < n; ++j) {
        const int m = n - j - 1;
        hipLaunchKernelGGL(k_house, dim3(1), dim3(kT), 0, s, a, n, j, d, e, vn);
        hipLaunchKernelGGL(k_symv, dim3(ceil_div(m, kWaves)), dim3(kT), 0, s, a, n, j,
                           (const double*)vn, p);
        hipLaunchKernelGGL(k_rank2, dim3(ceil_div(m, kRank2Rows)), dim3(kT), 0, s, a, n, j,
                           (const double*)vn, (const double*)p);
    }
    SI_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tail, dim3(1), dim3(64), 0, s, a, n, d, e);
    hipLaunchKernelGGL(k_tri_prep, dim3(1), dim3(kT), 0, s, d, e, n, e2, bounds);
    hipLaunchKernelGGL(k_bisect, dim3(k), dim3(kT), 0, s, d, e2, n, (const double*)bounds, w);
    hipLaunchKernelGGL(k_invit, dim3(1), dim3(kT), 0, s, d, e, n, k, (const double*)w,
                       (const double*)bounds, lu, piv, y, (uint64_t)0x5eedULL);
    hipLaunchKernelGGL(k_backtransform, dim3(k), dim3(kT), 0, s, (const double*)a, n,
                       (const double*)vn, y, evecs_dev, k);
    SI_LAUNCH_CHECK();
    to_host(evals_host, w, k, s);
}

}  // namespace si
