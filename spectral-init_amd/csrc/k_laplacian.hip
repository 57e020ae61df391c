// k_laplacian.hip -- device build of the symmetric normalised Laplacian
// (reference: src/laplacian.rs:8-88).
//
// Degrees.  PythonCompat accumulates deg_f32[col] += w over the nonzeros in
// CSR order (laplacian.rs:8-16).  For a bitwise-symmetric W with ascending
// columns, column j's entries in row order are row j's entries in column
// order, so one thread summing row j sequentially in f32 reproduces the
// scatter bit for bit.  The build verifies symmetry on the device; an
// asymmetric input falls back to an f32 atomic scatter (sum order not fixed).
// RustNative sums each row in f64 (laplacian.rs:41-46).
//
// Values.  L_ij = -isd_i * w_ij * isd_j evaluated left to right in f64
// (laplacian.rs:70-80); W's own diagonal is dropped and L's unit diagonal is
// implied (laplacian.rs:83-85).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <utility>
#include <vector>

#include "si_host.hpp"

namespace si {

namespace {

enum Flag { kFlagIndex = 0, kFlagWeight = 1, kFlagUnsorted = 2, kFlagAsym = 3, kNumFlags = 4 };

__global__ void k_convert_rowptr(const int64_t* __restrict__ ip64, int32_t* __restrict__ ip32,
                                 int64_t n1) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n1) ip32[i] = (int32_t)ip64[i];
}

// One thread per row: bounds, finiteness / sign of weights, ascending columns.
// The row kernels give kRowLanes lanes to a row; a workgroup takes
// 256 / kRowLanes consecutive rows at a time, grid-strided over at most
// kRowGrid workgroups.
constexpr int kRowLanes = 16;
constexpr int64_t kRowsPerBlock = 256 / kRowLanes;
constexpr int64_t kRowGrid = 4096;

inline dim3 row_grid(int64_t n) {
    return dim3((unsigned)std::min<int64_t>(ceil_div(n, kRowsPerBlock), kRowGrid));
}

// Row pointers and column indices in range, weights finite and non-negative,
// columns strictly increasing per row (else the unsorted flag); the longest
// row goes to *max_len (one atomic per workgroup).
__global__ __launch_bounds__(256) void k_validate(const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ col,
                                                  const float* __restrict__ w, int64_t n,
                                                  int* __restrict__ flags,
                                                  int32_t* __restrict__ max_len) {
    __shared__ int32_t wmax[4];
    const int sub = threadIdx.x & (kRowLanes - 1);
    int32_t longest = 0;
    for (int64_t base = (int64_t)blockIdx.x * kRowsPerBlock; base < n;
         base += (int64_t)gridDim.x * kRowsPerBlock) {
        const int64_t i = base + threadIdx.x / kRowLanes;
        if (i >= n) continue;
        const int32_t b = rp[i], e = rp[i + 1];
        if (e < b) {
            if (sub == 0) atomicOr(&flags[kFlagIndex], 1);
            continue;
        }
        longest = max(longest, e - b);
        for (int32_t p = b + sub; p < e; p += kRowLanes) {
            const int32_t j = col[p];
            if (j < 0 || j >= n) {
                atomicOr(&flags[kFlagIndex], 1);
                break;
            }
            const float v = w[p];
            if (!(v >= 0.0f) || isinf(v)) atomicOr(&flags[kFlagWeight], 1);
            if (p > b && j <= col[p - 1]) atomicOr(&flags[kFlagUnsorted], 1);
        }
    }
    for (int o = 32; o > 0; o >>= 1) longest = max(longest, __shfl_xor(longest, o));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = longest;
    __syncthreads();
    if (threadIdx.x == 0) {
        longest = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
        if (longest > 0) atomicMax(max_len, longest);
    }
}

// Each entry (i, j, w) must have a bitwise-equal (j, i, w).  Rows are strictly
// increasing here (no repeated columns), so it suffices that every entry
// above the diagonal finds its mirror and that the entries above and below
// the diagonal are equally many: the mirror map from the upper entries into
// the lower ones is then one-to-one and onto.  tri[0] / tri[1] count the
// upper / lower entries (one atomic each per workgroup).
__global__ __launch_bounds__(256) void k_check_symmetric(const int32_t* __restrict__ rp,
                                                         const int32_t* __restrict__ col,
                                                         const float* __restrict__ w, int64_t n,
                                                         int* __restrict__ flags,
                                                         int* __restrict__ tri) {
    __shared__ int wsum[2][4];
    const int sub = threadIdx.x & (kRowLanes - 1);
    int upper = 0, lower = 0;
    bool bad = false;
    for (int64_t base = (int64_t)blockIdx.x * kRowsPerBlock; base < n && !bad;
         base += (int64_t)gridDim.x * kRowsPerBlock) {
        const int64_t i = base + threadIdx.x / kRowLanes;
        if (i >= n) continue;
        const int32_t e = rp[i + 1];
        for (int32_t p = rp[i] + sub; p < e; p += kRowLanes) {
            const int32_t j = col[p];
            if (j < i) {
                ++lower;
                continue;
            }
            if (j == i) continue;
            ++upper;
            int32_t lo = rp[j], hi = rp[j + 1] - 1;
            bool found = false;
            while (lo <= hi) {
                const int32_t mid = (lo + hi) >> 1;
                const int32_t c = col[mid];
                if (c == i) {
                    found = __float_as_uint(w[mid]) == __float_as_uint(w[p]);
                    break;
                }
                if (c < i) lo = mid + 1;
                else hi = mid - 1;
            }
            if (!found) {
                bad = true;
                break;
            }
        }
    }
    if (bad) atomicOr(&flags[kFlagAsym], 1);
    for (int o = 32; o > 0; o >>= 1) {
        upper += __shfl_xor(upper, o);
        lower += __shfl_xor(lower, o);
    }
    if ((threadIdx.x & 63) == 0) {
        wsum[0][threadIdx.x >> 6] = upper;
        wsum[1][threadIdx.x >> 6] = lower;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&tri[0], wsum[0][0] + wsum[0][1] + wsum[0][2] + wsum[0][3]);
        atomicAdd(&tri[1], wsum[1][0] + wsum[1][1] + wsum[1][2] + wsum[1][3]);
    }
}

__global__ void k_degree_rowsum_f32(const int32_t* __restrict__ rp, const float* __restrict__ w,
                                    int64_t n, double* __restrict__ deg) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = 0.0f;
    for (int32_t p = rp[i]; p < rp[i + 1]; ++p) acc += w[p];
    deg[i] = (double)acc;
}

__global__ void k_degree_colscatter_f32(const int32_t* __restrict__ rp,
                                        const int32_t* __restrict__ col,
                                        const float* __restrict__ w, int64_t n,
                                        float* __restrict__ acc) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int32_t p = rp[i]; p < rp[i + 1]; ++p) atomicAdd(&acc[col[p]], w[p]);
}

__global__ void k_widen(const float* __restrict__ a, double* __restrict__ b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

__global__ void k_degree_rowsum_f64(const int32_t* __restrict__ rp, const float* __restrict__ w,
                                    int64_t n, double* __restrict__ deg) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int32_t p = rp[i]; p < rp[i + 1]; ++p) acc += (double)w[p];
    deg[i] = acc;
}

__global__ void k_isd(const double* __restrict__ deg, int64_t n, double* __restrict__ sq,
                      double* __restrict__ isd) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = sqrt(deg[i]);
    sq[i] = s;
    isd[i] = s > 0.0 ? 1.0 / s : 0.0;  // lib.rs:233-236
}

__global__ __launch_bounds__(256) void k_values(const int32_t* __restrict__ rp,
                                                const int32_t* __restrict__ col,
                                                const float* __restrict__ w,
                                                const double* __restrict__ isd, int64_t n,
                                                double* __restrict__ l64, float* __restrict__ a32,
                                                int32_t* __restrict__ long_rows,
                                                int* __restrict__ n_long) {
    const int sub = threadIdx.x & (kRowLanes - 1);
    for (int64_t base = (int64_t)blockIdx.x * kRowsPerBlock; base < n;
         base += (int64_t)gridDim.x * kRowsPerBlock) {
        const int64_t i = base + threadIdx.x / kRowLanes;
        if (i >= n) continue;
        const double di = isd[i];
        const int32_t b = rp[i], e = rp[i + 1];
        for (int32_t p = b + sub; p < e; p += kRowLanes) {
            const int32_t j = col[p];
            const double v = (j == i) ? 0.0 : (-di * (double)w[p]) * isd[j];
            l64[p] = v;
            a32[p] = (float)(-v);
        }
        if (sub == 0 && e - b > kLongRow) {
            const int slot = atomicAdd(n_long, 1);
            long_rows[slot] = (int32_t)i;
        }
    }
}

}  // namespace

void laplacian_free(DevLaplacian& L) {
    hipFree(L.rowptr); hipFree(L.col); hipFree(L.a32); hipFree(L.l64);
    hipFree(L.deg); hipFree(L.sqrt_deg); hipFree(L.isd); hipFree(L.long_rows);
    hipFree(L.perm); hipFree(L.iperm); hipFree(L.comp);
    L = DevLaplacian{};
}

void laplacian_build_device(DevLaplacian& L, int64_t n, const int64_t* d_indptr,
                            const int32_t* d_indices, const float* d_w, int64_t nnz, int mode,
                            hipStream_t s, bool reorder, const int64_t* h_indptr,
                            const int32_t* h_indices, bool prefer_rcm) {
    SI_CHECK(n > 0, kErrInvalidGraph, "laplacian: n must be > 0");
    SI_CHECK(nnz < (int64_t)INT32_MAX && n < (int64_t)INT32_MAX, kErrUnsupported,
             "laplacian: more than 2^31-1 stored entries or nodes is not supported");
    laplacian_free(L);
    L.n = n;
    L.nnz = nnz;
    L.degree_mode = mode;
    auto tp = std::chrono::steady_clock::now();
    double alloc0 = devbuf_alloc_ms();
    auto phase = [&](const char* what) {  // SI_TRACE=1: build phase timings
        if (!trace_on()) return;
        SI_HIP(hipStreamSynchronize(s));
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[build] %-24s %8.1f ms (device allocations %.1f ms)\n", what,
                     std::chrono::duration<double, std::milli>(now - tp).count(),
                     devbuf_alloc_ms() - alloc0);
        alloc0 = devbuf_alloc_ms();
        tp = now;
    };
    const int T = 256;
    const int64_t gn = ceil_div(n, T);
    SI_HIP(hipMalloc(&L.rowptr, sizeof(int32_t) * (n + 1)));
    SI_HIP(hipMalloc(&L.col, sizeof(int32_t) * (nnz > 0 ? nnz : 1)));
    SI_HIP(hipMalloc(&L.a32, sizeof(float) * (nnz > 0 ? nnz : 1)));
    SI_HIP(hipMalloc(&L.l64, sizeof(double) * (nnz > 0 ? nnz : 1)));
    SI_HIP(hipMalloc(&L.deg, sizeof(double) * n));
    SI_HIP(hipMalloc(&L.sqrt_deg, sizeof(double) * n));
    SI_HIP(hipMalloc(&L.isd, sizeof(double) * n));
    SI_HIP(hipMalloc(&L.long_rows, sizeof(int32_t) * (n > 0 ? n : 1)));
    int* d_flags = nullptr;
    SI_HIP(hipMalloc(&d_flags, sizeof(int) * (kNumFlags + 4)));
    SI_HIP(hipMemsetAsync(d_flags, 0, sizeof(int) * (kNumFlags + 4), s));
    int32_t* d_maxlen = d_flags + kNumFlags;
    int* d_nlong = d_flags + kNumFlags + 1;
    int* d_tri = d_flags + kNumFlags + 2;  // entries above / below the diagonal

    hipLaunchKernelGGL(k_convert_rowptr, dim3(ceil_div(n + 1, T)), dim3(T), 0, s, d_indptr,
                       L.rowptr, n + 1);
    if (nnz > 0) SI_HIP(hipMemcpyAsync(L.col, d_indices, sizeof(int32_t) * nnz,
                                       hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_validate, row_grid(n), dim3(T), 0, s, L.rowptr, L.col, d_w, n, d_flags,
                       d_maxlen);
    SI_LAUNCH_CHECK();
    int flags[kNumFlags + 4];
    SI_HIP(hipMemcpyAsync(flags, d_flags, sizeof(flags), hipMemcpyDeviceToHost, s));
    SI_HIP(hipStreamSynchronize(s));
    if (flags[kFlagIndex]) {
        hipFree(d_flags);
        laplacian_free(L);
        throw Error(kErrInvalidGraph, "invalid graph: row pointers or column indices out of range");
    }
    if (flags[kFlagWeight]) {
        hipFree(d_flags);
        laplacian_free(L);
        throw Error(kErrInvalidGraph, "invalid graph: weights must be finite and non-negative");
    }
    L.max_row_len = flags[kNumFlags];
    phase("validate");
    int sorted = flags[kFlagUnsorted] == 0;
    if (sorted) {
        hipLaunchKernelGGL(k_check_symmetric, row_grid(n), dim3(T), 0, s, L.rowptr, L.col, d_w, n,
                           d_flags, d_tri);
        SI_LAUNCH_CHECK();
        SI_HIP(hipMemcpyAsync(flags, d_flags, sizeof(flags), hipMemcpyDeviceToHost, s));
        SI_HIP(hipStreamSynchronize(s));
    }
    L.symmetric = sorted && flags[kFlagAsym] == 0 && flags[kNumFlags + 2] == flags[kNumFlags + 3];
    phase("symmetry check");

    if (mode == 0) {
        if (L.symmetric) {
            hipLaunchKernelGGL(k_degree_rowsum_f32, dim3(gn), dim3(T), 0, s, L.rowptr, d_w, n,
                               L.deg);
            L.degree_exact = 1;
        } else {
            float* acc = nullptr;
            SI_HIP(hipMalloc(&acc, sizeof(float) * n));
            SI_HIP(hipMemsetAsync(acc, 0, sizeof(float) * n, s));
            hipLaunchKernelGGL(k_degree_colscatter_f32, dim3(gn), dim3(T), 0, s, L.rowptr, L.col,
                               d_w, n, acc);
            hipLaunchKernelGGL(k_widen, dim3(gn), dim3(T), 0, s, acc, L.deg, n);
            SI_HIP(hipStreamSynchronize(s));
            hipFree(acc);
            L.degree_exact = 0;
        }
    } else {
        hipLaunchKernelGGL(k_degree_rowsum_f64, dim3(gn), dim3(T), 0, s, L.rowptr, d_w, n, L.deg);
        L.degree_exact = 1;
    }
    hipLaunchKernelGGL(k_isd, dim3(gn), dim3(T), 0, s, L.deg, n, L.sqrt_deg, L.isd);
    SI_LAUNCH_CHECK();
    phase("degrees");
    const float* w_vals = d_w;
    float* d_w2 = nullptr;
    StreamBuf<float> w2buf;  // the weights in the new order
    if (reorder && n > 1) {
        w2buf.alloc(nnz > 0 ? nnz : 1, s);
        d_w2 = w2buf.get();
        SI_HIP(hipMalloc(&L.perm, sizeof(int32_t) * n));
        SI_HIP(hipMalloc(&L.iperm, sizeof(int32_t) * n));
        if (L.symmetric) {
            SI_HIP(hipMalloc(&L.comp, sizeof(int32_t) * n));
            StreamBuf<int32_t> root(n, s);
            L.n_comp = components_device(n, d_indptr, d_indices, L.comp, root.get(), s);
            phase("components (device)");
            // k-NN-like graphs only: a mesh-like graph (average degree below
            // kClusterOrderMinDegree) gets a narrow band from RCM, which both
            // the SpMM and the banded Cholesky of level 2 use
            if (L.n_comp == 1 && n >= kClusterOrderMin && nnz >= kClusterOrderMinDegree * n &&
                !prefer_rcm) {
                // the clustering order walks on the operator in the input order:
                // its values first (rebuilt below in the new order)
                hipLaunchKernelGGL(k_values, row_grid(n), dim3(T), 0, s, L.rowptr, L.col, d_w, L.isd,
                                   n, L.l64, L.a32, L.long_rows, d_nlong);
                SI_LAUNCH_CHECK();
                int nl = 0;
                SI_HIP(hipMemcpyAsync(&nl, d_nlong, sizeof(int), hipMemcpyDeviceToHost, s));
                SI_HIP(hipStreamSynchronize(s));
                SI_HIP(hipMemsetAsync(d_nlong, 0, sizeof(int), s));
                L.n_long = nl;
                cluster_order_device(L, L.perm, L.iperm, s);
                L.order = 2;
                phase("cluster order (device)");
            } else {
                rcm_order_labeled(n, d_indptr, d_indices, L.comp, root.get(), L.n_comp, L.perm,
                                  L.iperm, s);
                L.order = 1;
                phase("rcm order (device)");
            }
        } else {
            // an asymmetric pattern: the host order, whose sweeps follow the
            // stored (out-)edges; pattern downloaded if the caller has no copy
            std::vector<int64_t> hp_buf;
            std::vector<int32_t> hx_buf;
            const int64_t* hp = h_indptr;
            const int32_t* hx = h_indices;
            if (!hp || !hx) {
                hp_buf.resize(n + 1);
                hx_buf.resize(nnz);
                SI_HIP(hipMemcpyAsync(hp_buf.data(), d_indptr, sizeof(int64_t) * (n + 1),
                                      hipMemcpyDeviceToHost, s));
                if (nnz)
                    SI_HIP(hipMemcpyAsync(hx_buf.data(), d_indices, sizeof(int32_t) * nnz,
                                          hipMemcpyDeviceToHost, s));
                SI_HIP(hipStreamSynchronize(s));
                hp = hp_buf.data();
                hx = hx_buf.data();
            }
            std::vector<int32_t> perm = rcm_order(n, hp, hx);
            std::vector<int32_t> iperm(n);
            for (int64_t i = 0; i < n; ++i) iperm[perm[i]] = (int32_t)i;
            SI_HIP(hipMemcpyAsync(L.perm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
            SI_HIP(hipMemcpyAsync(L.iperm, iperm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
            L.order = 1;
            phase("rcm order (host)");
        }
        // rows renumbered and sorted on the device, from the input CSR
        permute_csr_device(n, nnz, d_indptr, d_indices, d_w, L.perm, L.iperm, L.rowptr, L.col,
                           d_w2, s);
        phase("permuted csr");
        // node vectors follow the permutation (computed in input order above)
        StreamBuf<double> tmp(n, s);
        for (double* vec : {L.deg, L.sqrt_deg, L.isd}) {
            gather_rows_f64(vec, 1, tmp.get(), 1, n, 1, L.perm, s);
            SI_HIP(hipMemcpyAsync(vec, tmp.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        }
        SI_HIP(hipStreamSynchronize(s));
        w_vals = d_w2;
        phase("upload permuted");
    }
    hipLaunchKernelGGL(k_values, row_grid(n), dim3(T), 0, s, L.rowptr, L.col, w_vals, L.isd, n, L.l64,
                       L.a32, L.long_rows, d_nlong);
    SI_LAUNCH_CHECK();
    int nlong = 0;
    SI_HIP(hipMemcpyAsync(&nlong, d_nlong, sizeof(int), hipMemcpyDeviceToHost, s));
    SI_HIP(hipStreamSynchronize(s));
    L.n_long = nlong;
    phase("values");
    hipFree(d_flags);
}

// A Laplacian given by its FP64 off-diagonal values (unit diagonal implied),
// e.g. the centroid graph of the spectral meta-embedding, whose Gaussian
// affinities (often ~1e-80) exist only in FP64 before normalisation
// (src/multi_component.rs:170-205).  The FP32 sweep values are the rounded
// normalised entries; sqrt_deg (the trivial-vector hint) is stored scaled to
// max 1 so it survives FP32.
void laplacian_from_values_device(DevLaplacian& L, int64_t n, const int64_t* h_ip,
                                  const int32_t* h_ix, const double* h_l, const double* h_deg,
                                  hipStream_t s) {
    SI_CHECK(n > 0 && h_ip && h_ip[0] == 0, kErrInvalidArgument, "laplacian_from_values: bad CSR");
    const int64_t nnz = h_ip[n];
    SI_CHECK(nnz < (int64_t)INT32_MAX, kErrUnsupported, "laplacian_from_values: nnz >= 2^31");
    L = DevLaplacian{};
    L.n = n;
    L.nnz = nnz;
    L.degree_mode = 1;
    L.symmetric = 1;
    L.degree_exact = 1;
    std::vector<int32_t> rp(n + 1), lr;
    std::vector<float> a32(nnz > 0 ? nnz : 1);
    std::vector<double> sq(n), isd(n);
    int32_t maxlen = 0;
    double sqmax = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        rp[i] = (int32_t)h_ip[i];
        const int32_t len = (int32_t)(h_ip[i + 1] - h_ip[i]);
        maxlen = std::max(maxlen, len);
        if (len > kLongRow) lr.push_back((int32_t)i);
        for (int64_t p = h_ip[i]; p < h_ip[i + 1]; ++p) {
            SI_CHECK(h_ix[p] >= 0 && h_ix[p] < n && h_ix[p] != i, kErrInvalidArgument,
                     "laplacian_from_values: column out of range or on the diagonal");
            a32[p] = (float)(-h_l[p]);
        }
        sq[i] = std::sqrt(std::max(h_deg[i], 0.0));
        isd[i] = h_deg[i] > 0.0 ? 1.0 / sq[i] : 0.0;
        sqmax = std::max(sqmax, sq[i]);
    }
    rp[n] = (int32_t)nnz;
    if (sqmax > 0.0)
        for (double& v : sq) v /= sqmax;
    L.max_row_len = maxlen;
    L.n_long = (int64_t)lr.size();
    try {
        SI_HIP(hipMalloc(&L.rowptr, sizeof(int32_t) * (n + 1)));
        SI_HIP(hipMalloc(&L.col, sizeof(int32_t) * (nnz > 0 ? nnz : 1)));
        SI_HIP(hipMalloc(&L.a32, sizeof(float) * (nnz > 0 ? nnz : 1)));
        SI_HIP(hipMalloc(&L.l64, sizeof(double) * (nnz > 0 ? nnz : 1)));
        SI_HIP(hipMalloc(&L.deg, sizeof(double) * n));
        SI_HIP(hipMalloc(&L.sqrt_deg, sizeof(double) * n));
        SI_HIP(hipMalloc(&L.isd, sizeof(double) * n));
        SI_HIP(hipMalloc(&L.long_rows, sizeof(int32_t) * (lr.empty() ? 1 : lr.size())));
        SI_HIP(hipMemcpyAsync(L.rowptr, rp.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, s));
        if (nnz > 0) {
            SI_HIP(hipMemcpyAsync(L.col, h_ix, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, s));
            SI_HIP(hipMemcpyAsync(L.a32, a32.data(), sizeof(float) * nnz, hipMemcpyHostToDevice, s));
            SI_HIP(hipMemcpyAsync(L.l64, h_l, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
        }
        SI_HIP(hipMemcpyAsync(L.deg, h_deg, sizeof(double) * n, hipMemcpyHostToDevice, s));
        SI_HIP(hipMemcpyAsync(L.sqrt_deg, sq.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
        SI_HIP(hipMemcpyAsync(L.isd, isd.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
        if (!lr.empty())
            SI_HIP(hipMemcpyAsync(L.long_rows, lr.data(), sizeof(int32_t) * lr.size(),
                                  hipMemcpyHostToDevice, s));
        SI_HIP(hipStreamSynchronize(s));
    } catch (...) {
        laplacian_free(L);
        throw;
    }
}

}  // namespace si
