// k_dense.hip -- dense symmetric eigensolver for the k smallest eigenpairs of a
// Laplacian of up to kDenseMax nodes: the device form of dense_evd
// (src/solvers/dense.rs:9-46), used by escalation levels 0 and 5
// (src/solvers/mod.rs:83-109, 211-223) and by si_dense_evd.
//
//   1. densify L (FP64, row-major n x n, unit diagonal);
//   2. Householder tridiagonalisation, two launches per step: the next
//      reflector (k_tri_reflect), then the rank-2 update B -= v w^T + w v^T
//      (w = p - (v.p) v) of the trailing block fused with the next p = 2 B v
//      (k_tri_update), so the trailing block is read and written once per
//      step; reflectors are kept in R;
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

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
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

// One step j of the tridiagonalisation in two launches.  On entry the rows
// j+1.. of A hold B_j, the trailing block of step j (full symmetric storage);
// R row j holds the reflector v_j (vn[j] = 1) and p_in = 2 B_j v_j.
// k_tri_reflect (one workgroup): kk = v_j . p_in, the updated first row of B_j
// (matrix row j+1) and from it d[j+1] and the next reflector v_{j+1} into R
// row j+1 (vn[j+1], e[j+1]).  k_tri_update: every other row of B_j becomes a
// row of B_{j+1} = B_j - v w^T - w v^T (w = p_in - kk v_j), and its entry of
// p_out = 2 B_{j+1} v_{j+1} is formed while the row streams through, so the
// trailing block is read and written once per step.  j = -1 starts from the
// densified matrix (no update).
__global__ __launch_bounds__(kT) void k_tri_reflect(const double* __restrict__ A,
                                                    double* __restrict__ R, int n, int j,
                                                    double* __restrict__ d,
                                                    double* __restrict__ e,
                                                    double* __restrict__ vn,
                                                    const double* __restrict__ p_in,
                                                    double* __restrict__ scal) {
    __shared__ double red[kWaves];
    const int t = threadIdx.x;
    const int m = n - j - 1;  // order of B_j (>= 2)
    const bool refl = j >= 0 && vn[j] != 0.0;
    const double* vj = refl ? R + (int64_t)j * n + (j + 1) : nullptr;
    double kk = 0.0;
    if (refl) {
        double s = 0.0;
#pragma unroll 4
        for (int q = t; q < m; q += kT) s += vj[q] * p_in[q];
        kk = block_sum(s, red);
    }
    const double* b0 = A + (int64_t)(j + 1) * n + (j + 1);  // B_j row 0
    const double v0 = refl ? vj[0] : 0.0;
    const double w0 = refl ? p_in[0] - kk * v0 : 0.0;
    auto nr = [&](int q) -> double {  // updated B_j row 0, entry q
        double val = b0[q];
        if (refl) val -= v0 * (p_in[q] - kk * vj[q]) + w0 * vj[q];
        return val;
    };
    const double x0 = nr(1);
    double s2 = 0.0;
#pragma unroll 4
    for (int q = 2 + t; q < m; q += kT) {
        const double xq = nr(q);
        s2 += xq * xq;
    }
    const double tail2 = block_sum(s2, red);
    const bool refl1 = tail2 != 0.0;
    double alpha = x0;
    if (refl1) {
        const double nx = sqrt(x0 * x0 + tail2);
        alpha = x0 >= 0.0 ? -nx : nx;
        const double vv0 = x0 - alpha;
        const double inv = 1.0 / sqrt(vv0 * vv0 + tail2);
        double* r1 = R + (int64_t)(j + 1) * n + (j + 2);
        for (int q = t; q < m - 1; q += kT) r1[q] = (q == 0 ? vv0 : nr(q + 1)) * inv;
    }
    if (t == 0) {
        d[j + 1] = nr(0);
        e[j + 1] = alpha;
        vn[j + 1] = refl1 ? 1.0 : 0.0;
        scal[0] = kk;
    }
}

constexpr int kRowsPerWave = 1;
constexpr int kStepRows = kRowsPerWave * kWaves;  // rows of B_j per workgroup
__global__ __launch_bounds__(kT) void k_tri_update(double* __restrict__ A,
                                                   const double* __restrict__ R, int n, int j,
                                                   const double* __restrict__ vn,
                                                   const double* __restrict__ p_in,
                                                   const double* __restrict__ scal,
                                                   double* __restrict__ p_out) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int m = n - j - 1;
    const bool refl = j >= 0 && vn[j] != 0.0;
    const bool refl1 = vn[j + 1] != 0.0;
    const double kk = scal[0];
    const double* vj = R + (int64_t)(j >= 0 ? j : 0) * n + (j + 1);
    const double* vnext = R + (int64_t)(j + 1) * n + (j + 2);
    // rows r >= 1 of B_j become rows r - 1 of B_{j+1}; a wavefront streams its
    // rows together (independent loads in flight, v_j and p_in loaded once)
    const int rbase = blockIdx.x * kStepRows + wv * kRowsPerWave;
    double vr[kRowsPerWave], wr[kRowsPerWave], acc[kRowsPerWave];
    double* rows[kRowsPerWave];
    bool act[kRowsPerWave];
#pragma unroll
    for (int u = 0; u < kRowsPerWave; ++u) {
        const int r = rbase + u;
        act[u] = r >= 1 && r < m;
        rows[u] = A + (int64_t)(j + 1 + (act[u] ? r : 1)) * n + (j + 1);
        vr[u] = (refl && act[u]) ? vj[r] : 0.0;
        wr[u] = (refl && act[u]) ? p_in[r] - kk * vr[u] : 0.0;
        acc[u] = 0.0;
    }
#pragma unroll 4
    for (int q = 1 + lane; q < m; q += 64) {
        const double vq = refl ? vj[q] : 0.0;
        const double wq = refl ? p_in[q] - kk * vq : 0.0;
        const double vnq = refl1 ? vnext[q - 1] : 0.0;
#pragma unroll
        for (int u = 0; u < kRowsPerWave; ++u) {
            if (!act[u]) continue;
            double val = rows[u][q];
            if (refl) {
                val -= vr[u] * wq + wr[u] * vq;
                rows[u][q] = val;
            }
            acc[u] += val * vnq;
        }
    }
#pragma unroll
    for (int u = 0; u < kRowsPerWave; ++u) {
        const double a = wave_sum(acc[u]);
        if (lane == 0 && act[u]) p_out[rbase + u - 1] = refl1 ? 2.0 * a : 0.0;
    }
}

// The remaining steps j0 .. n-3 once the trailing block has at most kFinish
// rows, in one workgroup with the block in LDS: the same arithmetic as
// k_tri_reflect + k_tri_update (rank-2 update, next reflector, next p) with
// barriers instead of launches.  Entry state as for step j0 (j0 = -1: the whole
// matrix, no pending reflector); also writes d[n-1].
constexpr int kFinish = 128;
__global__ __launch_bounds__(kT) void k_tri_finish(const double* __restrict__ A,
                                                   double* __restrict__ R, int n, int j0,
                                                   double* __restrict__ d, double* __restrict__ e,
                                                   double* __restrict__ vn,
                                                   const double* __restrict__ p_in) {
    __shared__ double Bs[kFinish][kFinish + 1];
    __shared__ double vs[kFinish], ps[kFinish], ws[kFinish];
    __shared__ double red[kWaves];
    const int t = threadIdx.x;
    const int m0 = n - j0 - 1;
    for (int x = t; x < m0 * m0; x += kT) {
        const int r = x / m0, q = x % m0;
        Bs[r][q] = A[(int64_t)(j0 + 1 + r) * n + (j0 + 1 + q)];
    }
    bool refl = j0 >= 0 && vn[j0] != 0.0;
    if (refl)
        for (int q = t; q < m0; q += kT) {
            vs[q] = R[(int64_t)j0 * n + (j0 + 1 + q)];
            ps[q] = p_in[q];
        }
    __syncthreads();
    for (int j = j0; j + 2 < n; ++j) {
        const int o = j - j0;  // B_j = Bs[o.., o..]
        const int m = n - j - 1;
        if (refl) {
            double s = 0.0;
            for (int q = t; q < m; q += kT) s += vs[q] * ps[q];
            const double kk = block_sum(s, red);
            for (int q = t; q < m; q += kT) ws[q] = ps[q] - kk * vs[q];
            __syncthreads();
            for (int x = t; x < m * m; x += kT) {
                const int r = x / m, q = x % m;
                Bs[o + r][o + q] -= vs[r] * ws[q] + ws[r] * vs[q];
            }
            __syncthreads();
        }
        const double x0 = Bs[o][o + 1];
        double s2 = 0.0;
        for (int q = 2 + t; q < m; q += kT) s2 += Bs[o][o + q] * Bs[o][o + q];
        const double tail2 = block_sum(s2, red);
        const bool refl1 = tail2 != 0.0;
        double alpha = x0;
        if (refl1) {
            const double nx = sqrt(x0 * x0 + tail2);
            alpha = x0 >= 0.0 ? -nx : nx;
            const double vv0 = x0 - alpha;
            const double inv = 1.0 / sqrt(vv0 * vv0 + tail2);
            double* r1 = R + (int64_t)(j + 1) * n + (j + 2);
            for (int q = t; q < m - 1; q += kT) {
                const double vq = (q == 0 ? vv0 : Bs[o][o + 1 + q]) * inv;
                ws[q] = vq;
                r1[q] = vq;
            }
        }
        if (t == 0) {
            d[j + 1] = Bs[o][o];
            e[j + 1] = alpha;
            vn[j + 1] = refl1 ? 1.0 : 0.0;
        }
        __syncthreads();
        if (refl1) {
            for (int r = t; r < m - 1; r += kT) {
                double acc = 0.0;
                for (int q = 0; q < m - 1; ++q) acc += Bs[o + 1 + r][o + 1 + q] * ws[q];
                ps[r] = 2.0 * acc;
            }
            for (int q = t; q < m - 1; q += kT) vs[q] = ws[q];
        }
        __syncthreads();
        refl = refl1;
    }
    if (t == 0) d[n - 1] = Bs[n - 2 - j0][n - 2 - j0];
}

// the last diagonal entry once every step has run
__global__ void k_tri_last(const double* __restrict__ A, int n, double* __restrict__ d) {
    if (threadIdx.x == 0 && blockIdx.x == 0) d[n - 1] = A[(int64_t)(n - 1) * n + (n - 1)];
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
                    u0[i] = pv; u1[i] = cq; dl[i] = l; piv[i] = 0;
                    cp = dn - l * cq;
                    cq = en;
                } else {
                    const double l = cp / sub;
                    u0[i] = sub; u1[i] = dn; dl[i] = l; piv[i] = 1;
                    cp = cq - l * dn;
                    cq = -l * en;
                }
            }
            u0[n - 1] = fabs(cp) < pivmin ? (cp < 0.0 ? -pivmin : pivmin) : cp;
        }
        __syncthreads();
        for (int sweep = 0; sweep < 3; ++sweep) {
            if (t == 0) {
                // forward: row swaps and multipliers; the running row value stays
                // in a register, so no load waits on the previous store
                double c = yj[0];
                for (int i = 0; i + 1 < n; ++i) {
                    const double o = yj[i + 1];
                    const double l = dl[i];
                    double ri, ri1;
                    if (piv[i]) { ri = o; ri1 = c - l * o; }
                    else { ri = c; ri1 = o - l * c; }
                    yj[i] = ri;
                    c = ri1;
                }
                yj[n - 1] = c;
                // back substitution with the upper band (u0, u1, u2 = e[i+1] after a swap)
                double z1 = 0.0, z2 = 0.0, mx = 0.0;
                for (int i = n - 1; i >= 0; --i) {
                    const double u2 = (piv[i] && i + 2 < n) ? e[i + 1] : 0.0;
                    const double zi = (yj[i] - u1[i] * z1 - u2 * z2) / u0[i];
                    yj[i] = zi;
                    mx = fmax(mx, fabs(zi));
                    z2 = z1;
                    z1 = zi;
                }
                // scale down to keep later solves finite
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
    // working matrix A and the reflectors R (row j = v_j on columns j+1..)
    double* a = ws_get<double>(kWsDenseA, 2 * (size_t)n * n);
    double* rf = a + (size_t)n * n;
    // d, e, e2, vn, p (2 n, double-buffered) [6n] | lu [4n] | y [n k] | w [k] | bounds [4]
    double* aux = ws_get<double>(kWsDenseAux, (size_t)n * (10 + k) + k + 4);
    int* piv = ws_get<int>(kWsDensePiv, (size_t)n);
    double* d = aux;
    double* e = d + n;
    double* e2 = e + n;
    double* vn = e2 + n;
    double* p0 = vn + n;
    double* p1 = p0 + n;
    double* lu = p1 + n;
    double* y = lu + 4 * (size_t)n;
    double* w = y + (size_t)n * k;
    double* bounds = w + k;
    hipLaunchKernelGGL(k_densify, dim3(ceil_div(n, 256)), dim3(256), 0, s, L.rowptr, L.col, L.l64,
                       n, a);
    SI_LAUNCH_CHECK();
    double* pin = p1;
    double* pout = p0;
    int j = -1;
    for (; n - j - 1 > kFinish; ++j) {
        const int m = n - j - 1;
        hipLaunchKernelGGL(k_tri_reflect, dim3(1), dim3(kT), 0, s, (const double*)a, rf, n, j, d,
                           e, vn, (const double*)pin, bounds + 3);
        hipLaunchKernelGGL(k_tri_update, dim3(ceil_div(m, kStepRows)), dim3(kT), 0, s, a,
                           (const double*)rf, n, j, (const double*)vn, (const double*)pin,
                           (const double*)(bounds + 3), pout);
        std::swap(pin, pout);
    }
    SI_LAUNCH_CHECK();
    if (j + 2 < n) {
        hipLaunchKernelGGL(k_tri_finish, dim3(1), dim3(kT), 0, s, (const double*)a, rf, n, j, d, e,
                           vn, (const double*)pin);
    } else {
        hipLaunchKernelGGL(k_tri_last, dim3(1), dim3(64), 0, s, (const double*)a, n, d);
    }
    hipLaunchKernelGGL(k_tri_prep, dim3(1), dim3(kT), 0, s, d, e, n, e2, bounds);
    hipLaunchKernelGGL(k_bisect, dim3(k), dim3(kT), 0, s, d, e2, n, (const double*)bounds, w);
    hipLaunchKernelGGL(k_invit, dim3(1), dim3(kT), 0, s, d, e, n, k, (const double*)w,
                       (const double*)bounds, lu, piv, y, (uint64_t)0x5eedULL);
    hipLaunchKernelGGL(k_backtransform, dim3(k), dim3(kT), 0, s, (const double*)rf, n,
                       (const double*)vn, y, evecs_dev, k);
    SI_LAUNCH_CHECK();
    to_host(evals_host, w, k, s);
}

}  // namespace si
