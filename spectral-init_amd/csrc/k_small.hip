// k_small.hip -- small dense FP64 linear algebra in one workgroup, used for the
// Rayleigh-Ritz problems, Cholesky-QR factors and residual norms of the block
// eigensolvers (the roles of faer's SelfAdjointEigen and Cholesky at
// src/solvers/lobpcg.rs:158-187 and src/solvers/rsvd.rs:56-77).
//
// Eigenvalues use parallel cyclic Jacobi with a round-robin (circle) pairing:
// the n/2 disjoint rotations of a round are applied together, J^T A J one
// 2 x 2 block (two rows of one pair, two columns of another) per task, which
// is exact for disjoint pairs.  Matrices live in LDS with a padded row
// stride; n <= 64.
#include <type_traits>

#include "si_device.hpp"

namespace si {

namespace {

constexpr int N = kSmallMax;
constexpr int LDA = N + 1;
constexpr int kThreads = 256;
constexpr int kWide = 1024;

// Block size of the one-workgroup solvers.  The device code loops over
// blockDim.x, so any power of two up to kWide works; four wavefronts
// measured faster than one even at 16 rows (64 vs 72 us per generalized
// problem on MI355X), and the eigensolvers give sixteen to a matrix above 32
// rows, whose Jacobi rounds update up to 64 x 32 pairs of entries.
inline int small_threads(int /*n*/) { return kThreads; }

// Launch f(std::integral_constant<int, T>) with the eigensolvers' block size
// T for an n x n problem.
template <typename F>
void by_eig_threads(int n, F&& f) {
    if (n > 32) f(std::integral_constant<int, kWide>{});
    else f(std::integral_constant<int, kThreads>{});
}

// Pair (p, q) of round r for an even number of players m2 (circle method).
__device__ inline void round_pair(int r, int i, int m2, int& p, int& q) {
    const int m = m2 - 1;
    if (i == 0) {
        p = 0;
        q = 1 + (r % m);
    } else {
        p = 1 + (r + i) % m;
        q = 1 + (r - i + m) % m;
    }
    if (p > q) { int t = p; p = q; q = t; }
}

// Jacobi eigen-decomposition of the symmetric matrix in a (LDS, stride LDA);
// v receives the eigenvectors (columns).  On exit the diagonal of a holds the
// eigenvalues (unsorted).
__device__ void jacobi_lds(double (*a)[LDA], double (*v)[LDA], int n, double* cs, double* sn,
                           int* pp, int* qq, double* red) {
    const int t = threadIdx.x;
    const int m2 = n + (n & 1);
    const int half = m2 / 2;
    for (int e = t; e < n * n; e += (int)blockDim.x) v[e / n][e % n] = (e / n == e % n) ? 1.0 : 0.0;
    __syncthreads();
    for (int sweep = 0; sweep < 60; ++sweep) {
        // convergence test: off-diagonal mass relative to total (wavefront
        // butterflies, then the per-wave sums in a fixed order)
        double off = 0.0, tot = 0.0;
        for (int e = t; e < n * n; e += (int)blockDim.x) {
            const int i = e / n, j = e % n;
            const double q2 = a[i][j] * a[i][j];
            tot += q2;
            if (i != j) off += q2;
        }
        for (int o = 32; o > 0; o >>= 1) {
            off += __shfl_xor(off, o, 64);
            tot += __shfl_xor(tot, o, 64);
        }
        const int nw = ((int)blockDim.x + 63) / 64;
        if ((t & 63) == 0) { red[t >> 6] = off; red[nw + (t >> 6)] = tot; }
        __syncthreads();
        double offs = 0.0, tots = 0.0;
        for (int w = 0; w < nw; ++w) { offs += red[w]; tots += red[nw + w]; }
        __syncthreads();
        if (offs <= 1e-30 * tots || offs == 0.0 || n < 2) break;
        for (int r = 0; r < m2 - 1; ++r) {
            if (t < half) {
                int p, q;
                round_pair(r, t, m2, p, q);
                double c = 1.0, s = 0.0;
                if (q < n) {
                    const double apq = a[p][q];
                    if (apq != 0.0) {
                        const double app = a[p][p], aqq = a[q][q];
                        const double theta = (aqq - app) / (2.0 * apq);
                        const double tt = (theta >= 0.0 ? 1.0 : -1.0) /
                                          (fabs(theta) + sqrt(theta * theta + 1.0));
                        c = 1.0 / sqrt(tt * tt + 1.0);
                        s = tt * c;
                    }
                } else {
                    q = p;  // dummy player: identity
                }
                cs[t] = c; sn[t] = s; pp[t] = p; qq[t] = q;
            }
            __syncthreads();
            // A <- J^T A J one 2 x 2 block at a time: rows of pair P, columns
            // of pair Q (the blocks partition A); the annihilated entries of a
            // diagonal block are set to exactly zero.  A dummy pair (p == q)
            // has c = 1, s = 0 and writes its entries back unchanged.
            for (int e = t; e < half * half; e += (int)blockDim.x) {
                const int P = e / half, Q = e % half;
                const int p1 = pp[P], q1 = qq[P], p2 = pp[Q], q2 = qq[Q];
                const double c1 = cs[P], s1 = sn[P], c2 = cs[Q], s2 = sn[Q];
                const double a11 = a[p1][p2], a12 = a[p1][q2];
                const double a21 = a[q1][p2], a22 = a[q1][q2];
                const double u11 = c1 * a11 - s1 * a21, u12 = c1 * a12 - s1 * a22;
                const double u21 = s1 * a11 + c1 * a21, u22 = s1 * a12 + c1 * a22;
                const bool diag = P == Q && p1 != q1;
                a[p1][p2] = c2 * u11 - s2 * u12;
                a[p1][q2] = diag ? 0.0 : s2 * u11 + c2 * u12;
                a[q1][p2] = diag ? 0.0 : c2 * u21 - s2 * u22;
                a[q1][q2] = s2 * u21 + c2 * u22;
            }
            // V <- V J
            for (int e = t; e < half * n; e += (int)blockDim.x) {
                const int pi = e / n, j = e % n;
                const int p = pp[pi], q = qq[pi];
                if (p == q) continue;
                const double c = cs[pi], s = sn[pi];
                const double vjp = v[j][p], vjq = v[j][q];
                v[j][p] = c * vjp - s * vjq;
                v[j][q] = s * vjp + c * vjq;
            }
            __syncthreads();
        }
    }
}

// Sort the diagonal of a ascending and write eigenvalues w[n] and vectors
// out[n x n] (row-major, column j = vector j) from the columns of v.
__device__ void sort_and_store(double (*a)[LDA], double (*v)[LDA], int n, int* order, double* w,
                               double* out) {
    const int t = threadIdx.x;
    // stable ascending order by rank: element i goes to the number of elements
    // below it plus the equal ones before it (NaN ranks as +inf)
    for (int i = t; i < n; i += (int)blockDim.x) {
        const double di = isnan(a[i][i]) ? INFINITY : a[i][i];
        int rank = 0;
        for (int j = 0; j < n; ++j) {
            const double dj = isnan(a[j][j]) ? INFINITY : a[j][j];
            rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
        }
        order[rank] = i;
    }
    __syncthreads();
    for (int i = t; i < n; i += (int)blockDim.x) w[i] = a[order[i]][order[i]];
    for (int e = t; e < n * n; e += (int)blockDim.x) {
        const int r = e / n, c = e % n;
        out[r * n + c] = v[r][order[c]];
    }
}

// In-place lower Cholesky of the LDS matrix; returns false (uniformly) if not SPD.
__device__ bool cholesky_lds(double (*l)[LDA], int n, int* flag) {
    const int t = threadIdx.x;
    if (t == 0) *flag = 0;
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        if (t == 0) {
            double s = l[j][j];
            for (int k = 0; k < j; ++k) s -= l[j][k] * l[j][k];
            if (!(s > 0.0)) *flag = 1;
            l[j][j] = s > 0.0 ? sqrt(s) : 1.0;
        }
        __syncthreads();
        const double d = l[j][j];
        for (int i = j + 1 + t; i < n; i += (int)blockDim.x) {
            double s = l[i][j];
            for (int k = 0; k < j; ++k) s -= l[i][k] * l[j][k];
            l[i][j] = s / d;
        }
        __syncthreads();
    }
    for (int e = t; e < n * n; e += (int)blockDim.x)
        if (e % n > e / n) l[e / n][e % n] = 0.0;
    __syncthreads();
    return *flag == 0;
}

template <int T>
__global__ __launch_bounds__(T) void k_sym_eig(int n, const double* __restrict__ ain,
                                                       int lda, double* __restrict__ w,
                                                       double* __restrict__ out) {
    __shared__ double a[N][LDA];
    __shared__ double v[N][LDA];
    __shared__ double cs[N / 2], sn[N / 2];
    __shared__ int pp[N / 2], qq[N / 2], order[N];
    __shared__ double red[2 * (T / 64)];
    const int t = threadIdx.x;
    for (int e = t; e < n * n; e += (int)blockDim.x) {
        const int i = e / n, j = e % n;
        // lower triangle defines the matrix (SelfAdjointEigen, Side::Lower)
        a[i][j] = (j <= i) ? ain[i * lda + j] : ain[j * lda + i];
    }
    __syncthreads();
    jacobi_lds(a, v, n, cs, sn, pp, qq, red);
    sort_and_store(a, v, n, order, w, out);
}

// A c = theta B c for strided inputs (lower triangles read).  With scale != 0
// the problem is first scaled by D = diag(B)^-1/2 on both sides (column
// normalisation of the underlying block, done on the small matrices) and the
// returned vectors are D c'.
// Batched over blockIdx.x with strides (in doubles) between problems: problem
// p reads ain + p*sa, bin + p*sb and writes w + p*n, cout + p*n*n, status[p].
template <int T>
__global__ __launch_bounds__(T) void k_gen_eig(int n, const double* __restrict__ ain,
                                                       int lda, const double* __restrict__ bin,
                                                       int ldb, int scale,
                                                       double* __restrict__ w,
                                                       double* __restrict__ cout,
                                                       int* __restrict__ status, int64_t sa,
                                                       int64_t sb, const int32_t* __restrict__ gate) {
    if (gate != nullptr && *gate == 0) return;  // device-planned round with nothing to do
    ain += blockIdx.x * sa;
    bin += blockIdx.x * sb;
    w += (int64_t)blockIdx.x * n;
    cout += (int64_t)blockIdx.x * n * n;
    status += blockIdx.x;
    __shared__ double a[N][LDA];
    __shared__ double l[N][LDA];
    __shared__ double v[N][LDA];
    __shared__ double cs[N / 2], sn[N / 2];
    __shared__ int pp[N / 2], qq[N / 2], order[N];
    __shared__ double red[2 * (T / 64)];
    __shared__ double dsc[N];
    __shared__ int flag;
    const int t = threadIdx.x;
    if (t < n) {
        const double bd = bin[t * ldb + t];
        dsc[t] = scale ? (bd > 0.0 ? 1.0 / sqrt(bd) : 0.0) : 1.0;
    }
    __syncthreads();
    for (int e = t; e < n * n; e += (int)blockDim.x) {
        const int i = e / n, j = e % n;
        const double s = dsc[i] * dsc[j];
        a[i][j] = s * ((j <= i) ? ain[i * lda + j] : ain[j * lda + i]);
        l[i][j] = s * ((j <= i) ? bin[i * ldb + j] : bin[j * ldb + i]);
    }
    __syncthreads();
    if (!cholesky_lds(l, n, &flag)) {
        if (t == 0) status[0] = 1;
        return;
    }
    // v = L^-1 A  (thread per column, forward substitution)
    for (int c = t; c < n; c += (int)blockDim.x)
        for (int i = 0; i < n; ++i) {
            double s = a[i][c];
            for (int k = 0; k < i; ++k) s -= l[i][k] * v[k][c];
            v[i][c] = s / l[i][i];
        }
    __syncthreads();
    // a = L^-1 v^T  (= L^-1 A L^-T), then symmetrise
    for (int c = t; c < n; c += (int)blockDim.x)
        for (int i = 0; i < n; ++i) {
            double s = v[c][i];
            for (int k = 0; k < i; ++k) s -= l[i][k] * a[k][c];
            a[i][c] = s / l[i][i];
        }
    __syncthreads();
    for (int e = t; e < n * n; e += (int)blockDim.x) {
        const int i = e / n, j = e % n;
        if (j < i) {
            const double m = 0.5 * (a[i][j] + a[j][i]);
            a[i][j] = m;
            a[j][i] = m;
        }
    }
    __syncthreads();
    jacobi_lds(a, v, n, cs, sn, pp, qq, red);
    // eigenvalues on the diagonal of a; c = L^-T v, written back into v
    for (int c = t; c < n; c += (int)blockDim.x)
        for (int i = n - 1; i >= 0; --i) {
            double s = v[i][c];
            for (int k = i + 1; k < n; ++k) s -= l[k][i] * v[k][c];
            v[i][c] = s / l[i][i];
        }
    __syncthreads();
    for (int e = t; e < n * n; e += (int)blockDim.x) v[e / n][e % n] *= dsc[e / n];
    __syncthreads();
    sort_and_store(a, v, n, order, w, cout);
    if (t == 0) status[0] = 0;
}

__global__ __launch_bounds__(kThreads) void k_chol_inv(int n, const double* __restrict__ g,
                                                        double* __restrict__ tout,
                                                        int* __restrict__ status) {
    __shared__ double l[N][LDA];
    __shared__ double ti[N][LDA];
    __shared__ int flag;
    const int t = threadIdx.x;
    for (int e = t; e < n * n; e += (int)blockDim.x) {
        const int i = e / n, j = e % n;
        l[i][j] = (j <= i) ? g[i * n + j] : g[j * n + i];
    }
    __syncthreads();
    if (!cholesky_lds(l, n, &flag)) {
        if (t == 0) status[0] = 1;
        return;
    }
    // T = L^-T: solve L^T T[:, c] = e_c (back substitution, thread per column)
    for (int c = t; c < n; c += (int)blockDim.x)
        for (int i = n - 1; i >= 0; --i) {
            double s = (i == c) ? 1.0 : 0.0;
            for (int k = i + 1; k < n; ++k) s -= l[k][i] * ti[k][c];
            ti[i][c] = s / l[i][i];
        }
    __syncthreads();
    for (int e = t; e < n * n; e += (int)blockDim.x) tout[e] = ti[e / n][e % n];
    if (t == 0) status[0] = 0;
}

__global__ void k_inv_sqrt_diag(int n, const double* __restrict__ g, double* __restrict__ sc) {
    const int c = threadIdx.x;
    if (c < n) {
        const double d = g[c * n + c];
        sc[c] = d > 0.0 ? 1.0 / sqrt(d) : 0.0;
    }
}

__global__ void k_residual_norms(int n, const double* __restrict__ gr, int ldr,
                                 const double* __restrict__ gx, int ldx, double* __restrict__ res) {
    const int c = threadIdx.x;
    if (c < n) {
        double xn = gx ? sqrt(fmax(gx[c * ldx + c], 0.0)) : 1.0;
        if (xn < 2.220446049250313e-16) xn = 2.220446049250313e-16;
        res[c] = sqrt(fmax(gr[c * ldr + c], 0.0)) / xn;
    }
}

__global__ void k_copy_sub(int rows, int cols, const double* __restrict__ src, int lds,
                           double* __restrict__ dst, int ldd, double alpha) {
    for (int e = threadIdx.x; e < rows * cols; e += blockDim.x) {
        const int r = e / cols, c = e % cols;
        dst[r * ldd + c] = alpha * src[r * lds + c];
    }
}

// SVQB transform (Stathopoulos & Wu, 2002): from G = U diag(w) U^T, the
// columns of U scaled by 1/sqrt(w) (zeroed where w <= rel * max(w)), so that
// Y T has orthonormal (or zero) columns.
__global__ void k_svqb(int n, const double* __restrict__ w, const double* __restrict__ u,
                       double* __restrict__ t, double rel) {
    double wmax = 0.0;
    for (int i = 0; i < n; ++i) wmax = fmax(wmax, w[i]);
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int c = e % n;
        const double wc = w[c];
        t[e] = (wc > rel * wmax && wc > 0.0) ? u[e] / sqrt(wc) : 0.0;
    }
}

// Dense EVD of the diagonal blocks of a block-diagonal L, one workgroup per
// block: block b covers rows [seg_row0[b], seg_row0[b] + seg_n[b]) (<= 64 rows,
// columns inside the same range).  The k smallest eigenvalues go to
// evals[b*k + j], their vectors to rows of evecs (leading dimension ldv).
template <int T>
__global__ __launch_bounds__(T) void k_dense_seg_eig(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
    const double* __restrict__ l64, const int64_t* __restrict__ seg_row0,
    const int32_t* __restrict__ seg_n, int k, double* __restrict__ evals,
    double* __restrict__ evecs, int ldv) {
    __shared__ double a[N][LDA];
    __shared__ double v[N][LDA];
    __shared__ double cs[N / 2], sn[N / 2];
    __shared__ int pp[N / 2], qq[N / 2], order[N];
    __shared__ double red[2 * (T / 64)];
    const int t = threadIdx.x;
    const int64_t r0 = seg_row0[blockIdx.x];
    const int n = seg_n[blockIdx.x];
    for (int e = t; e < n * n; e += (int)blockDim.x) a[e / n][e % n] = (e / n == e % n) ? 1.0 : 0.0;
    __syncthreads();
    for (int i = t; i < n; i += (int)blockDim.x)
        for (int32_t q = rp[r0 + i]; q < rp[r0 + i + 1]; ++q) {
            const int64_t j = (int64_t)col[q] - r0;
            if (j != i && j >= 0 && j < n) a[i][j] += l64[q];
        }
    __syncthreads();
    jacobi_lds(a, v, n, cs, sn, pp, qq, red);
    if (t == 0) {
        for (int i = 0; i < n; ++i) order[i] = i;
        for (int i = 1; i < n; ++i) {
            const int oi = order[i];
            int j = i - 1;
            while (j >= 0 && a[order[j]][order[j]] > a[oi][oi]) { order[j + 1] = order[j]; --j; }
            order[j + 1] = oi;
        }
    }
    __syncthreads();
    const int kk = k < n ? k : n;
    for (int j = t; j < kk; j += (int)blockDim.x) evals[(int64_t)blockIdx.x * k + j] = a[order[j]][order[j]];
    for (int e = t; e < n * kk; e += (int)blockDim.x) {
        const int r = e / kk, c = e % kk;
        evecs[(r0 + r) * ldv + c] = v[r][order[c]];
    }
}

__global__ void k_dense_from_csr(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                 const double* __restrict__ l64, int n, double* __restrict__ d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int j = 0; j < n; ++j) d[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int32_t p = rp[i]; p < rp[i + 1]; ++p)
        if (col[p] != i) d[i * n + col[p]] += l64[p];
}

}  // namespace

void small_svqb(int n, const double* w, const double* u, double* t, double rel, hipStream_t s) {
    hipLaunchKernelGGL(k_svqb, dim3(1), dim3(256), 0, s, n, w, u, t, rel);
    SI_LAUNCH_CHECK();
}

void dense_from_csr(const DevLaplacian& L, double* d, hipStream_t s) {
    SI_CHECK(L.n <= kSmallMax, kErrInvalidArgument, "dense_from_csr: n too large");
    hipLaunchKernelGGL(k_dense_from_csr, dim3(1), dim3(64), 0, s, L.rowptr, L.col, L.l64, (int)L.n, d);
    SI_LAUNCH_CHECK();
}

void small_copy_sub(int rows, int cols, const double* src, int lds, double* dst, int ldd,
                    double alpha, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_sub, dim3(1), dim3(256), 0, s, rows, cols, src, lds, dst, ldd, alpha);
    SI_LAUNCH_CHECK();
}

void small_sym_eig(int n, const double* a, int lda, double* w, double* v, hipStream_t s) {
    SI_CHECK(n >= 1 && n <= kSmallMax, kErrInvalidArgument, "small_sym_eig: n out of range");
    by_eig_threads(n, [&](auto t) {
        constexpr int T = decltype(t)::value;
        hipLaunchKernelGGL((k_sym_eig<T>), dim3(1), dim3(T), 0, s, n, a, lda, w, v);
    });
    SI_LAUNCH_CHECK();
}

void small_gen_eig(int n, const double* a, const double* b, double* w, double* c, int* status,
                   hipStream_t s) {
    SI_CHECK(n >= 1 && n <= kSmallMax, kErrInvalidArgument, "small_gen_eig: n out of range");
    by_eig_threads(n, [&](auto t) {
        constexpr int T = decltype(t)::value;
        hipLaunchKernelGGL((k_gen_eig<T>), dim3(1), dim3(T), 0, s, n, a, n, b, n, 0, w, c, status,
                           (int64_t)0, (int64_t)0, (const int32_t*)nullptr);
    });
    SI_LAUNCH_CHECK();
}

void small_gen_eig_scaled(int n, const double* a, int lda, const double* b, int ldb, double* w,
                          double* c, int* status, hipStream_t s) {
    SI_CHECK(n >= 1 && n <= kSmallMax, kErrInvalidArgument, "small_gen_eig: n out of range");
    by_eig_threads(n, [&](auto t) {
        constexpr int T = decltype(t)::value;
        hipLaunchKernelGGL((k_gen_eig<T>), dim3(1), dim3(T), 0, s, n, a, lda, b, ldb, 1, w, c, status,
                           (int64_t)0, (int64_t)0, (const int32_t*)nullptr);
    });
    SI_LAUNCH_CHECK();
}
void small_gen_eig_scaled_batched(int count, int n, const double* a, int lda, int64_t sa,
                                  const double* b, int ldb, int64_t sb, double* w, double* c,
                                  int* status, hipStream_t s, const int32_t* gate) {
    SI_CHECK(n >= 1 && n <= kSmallMax && count >= 1, kErrInvalidArgument,
             "small_gen_eig_batched: size out of range");
    by_eig_threads(n, [&](auto t) {
        constexpr int T = decltype(t)::value;
        hipLaunchKernelGGL((k_gen_eig<T>), dim3(count), dim3(T), 0, s, n, a, lda, b, ldb, 1, w, c,
                           status, sa, sb, gate);
    });
    SI_LAUNCH_CHECK();
}
void dense_seg_eig(const DevLaplacian& L, int count, const int64_t* seg_row0, const int32_t* seg_n,
                   int k, double* evals, double* evecs, int ldv, hipStream_t s) {
    if (count <= 0) return;
    // segments of up to 64 rows: the wide block
    hipLaunchKernelGGL((k_dense_seg_eig<kWide>), dim3(count), dim3(kWide), 0, s, L.rowptr, L.col,
                       L.l64, seg_row0, seg_n, k, evals, evecs, ldv);
    SI_LAUNCH_CHECK();
}

void small_chol_inv(int n, const double* g, double* t, int* status, hipStream_t s) {
    SI_CHECK(n >= 1 && n <= kSmallMax, kErrInvalidArgument, "small_chol_inv: n out of range");
    hipLaunchKernelGGL(k_chol_inv, dim3(1), dim3(small_threads(n)), 0, s, n, g, t, status);
    SI_LAUNCH_CHECK();
}

void small_inv_sqrt_diag(int n, const double* g, double* scale, hipStream_t s) {
    hipLaunchKernelGGL(k_inv_sqrt_diag, dim3(1), dim3(64), 0, s, n, g, scale);
    SI_LAUNCH_CHECK();
}

void small_residual_norms(int n, const double* gr, int ldr, const double* gx, int ldx,
                          double* res, hipStream_t s) {
    hipLaunchKernelGGL(k_residual_norms, dim3(1), dim3(64), 0, s, n, gr, ldr, gx, ldx, res);
    SI_LAUNCH_CHECK();
}

}  // namespace si
