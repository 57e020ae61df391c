// rsvd.cpp -- randomized SVD of M = 2I - L (src/solvers/rsvd.rs:135-213), the
// escalation chain's level 4 and the comparison solver of the 1M-node graph.
//
// M is applied through the SpMM epilogue (s_l = -1, s_x = 2): no second matrix
// is stored.  The sketch width is the reference's: rank + oversampling, with
// rsvd_k_sub (rsvd.rs:84-87) when no oversampling is given, up to n.
//
// The sketch block is stored with a padded leading dimension kp: the smallest
// width a single SpMM launch serves (1, 2, 3, 4, 8, 16, 32, 64; multiples of
// 64 beyond), the columns past ksub held at zero (M 0 = 0, so they stay zero).
// Every application of M is then one launch per 64 columns -- a 13-column
// sketch is one 16-wide launch, not 8 + 4 + 1.
//
// Orthonormal bases (the reference's thin Householder Q) come from SVQB
// applied twice: the Gram Y^T Y, its eigenpairs (w, U), Y <- Y U w^-1/2.  Both
// give an orthonormal basis of span(Y), which is all the later steps use (B =
// Q^T M Q and V = Q U_B are invariant to the basis).
//   ksub <= 64: the project's kernels -- FP64-MFMA Gram, one-workgroup Jacobi,
//               the 16-wide products on the FP64 matrix cores;
//   ksub > 64:  (the reference's n/10 oversampling for n > 610) rocBLAS DGEMM
//               for the Gram and block products, rocSOLVER DSYEVD for the
//               ksub x ksub eigenproblems.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>

#include "si_host.hpp"
#include "si_rocblas.hpp"

namespace si {

namespace {

__global__ void k_svqb_scale_cm(int k, const double* __restrict__ w, double rel, double* u) {
    // U[:, j] *= w_j^-1/2 (0 below rel * w_max); U column-major k x k
    const double wmax = w[k - 1];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)k * k;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(e / k);
        const double wj = w[j];
        u[e] *= wj > rel * wmax ? 1.0 / sqrt(wj) : 0.0;
    }
}

int sketch_storage_width(int64_t ksub) {
    const int widths[] = {1, 2, 3, 4, 8, 16, 32, 64};
    for (int w : widths)
        if (w >= ksub) return w;
    return (int)(ceil_div(ksub, (int64_t)64) * 64);
}

// Small sketches (ksub <= 64) on the project's kernels.  A sketch of 9-16
// columns is stored 16 wide with zero columns past ksub; its basis change
// then runs as the 16 x 16 product on the FP64 matrix cores, with the k x k
// transform zero-padded to 16 x 16 (the padded columns stay zero, the others
// are exactly Y T).
struct SmallOrth {
    DevBuf<double> g, w, u, t, t16, ws;
    SmallOrth(int64_t n, int k) : g((size_t)k * k), w(k), u((size_t)k * k), t((size_t)k * k),
                                  t16(256), ws(gram_workspace_doubles(n, k, k)) {}
    // y [n x ld], first k columns: replaced by an orthonormal basis of their span
    void orth(int64_t n, double* y, int ld, int k, hipStream_t s) {
        for (int pass = 0; pass < 2; ++pass) {
            gram_f64(n, y, ld, k, y, ld, k, g, ws, s);
            small_sym_eig(k, g, k, w, u, s);
            small_svqb(k, w, u, t, 1e-28, s);
            if (ld == 16) {
                SI_HIP(hipMemsetAsync(t16.get(), 0, sizeof(double) * 256, s));
                small_copy_sub(k, k, t, k, t16, 16, 1.0, s);
                block_mul_f64(n, y, 16, 16, t16, 16, nullptr, 0, 0, nullptr, y, 16, s);
            } else {
                block_mul_f64(n, y, ld, k, t, k, nullptr, 0, 0, nullptr, y, ld, s);
            }
        }
    }
};

// Large sketches (ksub > 64) on rocBLAS / rocSOLVER.  A row-major block Y
// [n x ld] is, to the column-major libraries, the ld x n matrix Y^T.
struct LargeOrth {
    int k;
    DevBuf<double> g, w, e, tmp;
    DevBuf<rocblas_int> info;
    LargeOrth(int64_t n, int k_, int ld) : k(k_), g((size_t)k_ * k_), w(k_), e(k_),
                                           tmp((size_t)n * ld), info(1) {
        // the product writes the first k columns only: the padding stays zero
        SI_HIP(hipMemset(tmp.get(), 0, sizeof(double) * (size_t)n * ld));
    }
    // g = A^T B for row-major A, B [n x ld] (first k columns): col-major k x k
    void gram(int64_t n, const double* a, const double* b, int ld, hipStream_t s) {
        const double one = 1.0, zero = 0.0;
        rb_check(rocblas_dgemm(rb_handle(s), rocblas_operation_none, rocblas_operation_transpose, k,
                               k, (rocblas_int)n, &one, a, ld, b, ld, &zero, g, k),
                 "rocblas_dgemm (Gram)");
    }
    // eigenpairs of g in place (ascending w; vectors in the columns of g)
    void eig(hipStream_t s) {
        rb_check(rocsolver_dsyevd(rb_handle(s), rocblas_evect_original, rocblas_fill_upper, k, g, k,
                                  w, e, info),
                 "rocsolver_dsyevd");
        rocblas_int hinfo = 0;
        to_host(&hinfo, info.get(), 1, s);
        SI_CHECK(hinfo == 0, kErrInternal, "rsvd: symmetric eigensolver did not converge");
    }
    // out[n x ld] (first q columns) = a[n x ld] (first k columns) * C, C
    // column-major k x q with leading dimension ldc
    void mul(int64_t n, const double* a, int ld, const double* c, int ldc, int q, double* out,
             int ldo, hipStream_t s) {
        const double one = 1.0, zero = 0.0;
        rb_check(rocblas_dgemm(rb_handle(s), rocblas_operation_transpose, rocblas_operation_none, q,
                               (rocblas_int)n, k, &one, c, ldc, a, ld, &zero, out, ldo),
                 "rocblas_dgemm (block product)");
    }
    void orth(int64_t n, double* y, int ld, hipStream_t s) {
        for (int pass = 0; pass < 2; ++pass) {
            gram(n, y, y, ld, s);
            eig(s);
            hipLaunchKernelGGL(k_svqb_scale_cm, dim3(256), dim3(256), 0, s, k, (const double*)w.get(),
                               1e-28, g.get());
            SI_LAUNCH_CHECK();
            mul(n, y, ld, g, k, k, tmp, ld, s);
            SI_HIP(hipMemcpyAsync(y, tmp.get(), sizeof(double) * (size_t)n * ld,
                                  hipMemcpyDeviceToDevice, s));
        }
    }
};

}  // namespace

void rsvd_solve_device(const DevLaplacian& L, int n_components, uint64_t seed, int oversampling,
                       int nbiter, double* evals_host, double* evecs_dev, SolveStats& st,
                       hipStream_t s) {
    const int64_t n = L.n;
    const int rank = n_components + 1;
    SI_CHECK(n_components > 0 && n_components < n, kErrInvalidArgument,
             "rsvd_solve: need 0 < n_components < n");
    SI_CHECK(nbiter >= 0, kErrInvalidArgument, "rsvd_solve: nbiter must be >= 0");
    int64_t ksub;
    if (oversampling < 0) {  // rsvd_k_sub (rsvd.rs:84-87)
        int64_t os = std::max<int64_t>(n / 10, std::max(rank, 5));
        os = std::min<int64_t>(os, n > rank ? n - rank : 0);
        ksub = rank + os;
    } else {
        ksub = rank + oversampling;
    }
    ksub = std::min<int64_t>(ksub, n);
    SI_CHECK(ksub >= rank, kErrUnsupported, "rsvd_solve: sketch narrower than the rank");
    const int k = (int)ksub;
    const int kp = sketch_storage_width(ksub);
    {
        size_t freeb = 0, totalb = 0;
        SI_HIP(hipMemGetInfo(&freeb, &totalb));
        const double need = 8.0 * (double)n * kp * (k > kSmallMax ? 3.0 : 2.0) + 16.0 * k * k;
        SI_CHECK(need < 0.8 * (double)freeb, kErrUnsupported,
                 "rsvd_solve: the " + std::to_string(k) + "-column sketch of " + std::to_string(n) +
                     " rows does not fit in device memory");
    }
    st = SolveStats{};
    st.block_size = k;
    const size_t nk = (size_t)n * kp;
    DevBuf<double> om(nk), y(nk);
    SI_HIP(hipMemsetAsync(om.get(), 0, sizeof(double) * nk, s));
    SI_HIP(hipMemsetAsync(y.get(), 0, sizeof(double) * nk, s));
    auto apply_m = [&](const double* in, double* out) {
        spmm_f64(L, in, out, kp, kp, -1.0, 2.0, 0.0, nullptr, s);  // (2I - L) x, all kp columns
        ++st.spmm_calls;
    };
    // the sketch: the same draws (index = original node * ksub + column) at any kp
    fill_normal_f64(om, n, k, kp, seed, kStreamRsvdOmega, s, L.perm);
    std::vector<double> hw(k);
    std::vector<int> cols(rank);
    cols[0] = k - 1;  // trivial pair = largest theta, then the next n_components (rsvd.rs:190-212)
    for (int c = 0; c < n_components; ++c) cols[c + 1] = std::max(0, k - 2 - c);
    if (k <= kSmallMax) {
        SmallOrth o(n, k);
        for (int it = 0; it < nbiter; ++it) {
            apply_m(om, y);
            o.orth(n, y, kp, k, s);
            apply_m(y, om);
            o.orth(n, om, kp, k, s);
        }
        apply_m(om, y);
        o.orth(n, y, kp, k, s);                         // Q
        apply_m(y, om);                                  // M Q
        gram_f64(n, y, kp, k, om, kp, k, o.g, o.ws, s);  // B = Q^T M Q
        small_sym_eig(k, o.g, k, o.w, o.u, s);           // ascending
        to_host(hw.data(), o.w.get(), k, s);
        // V = Q (U S): the eigenvector rotation and the column selection in one product
        std::vector<double> hu((size_t)k * k), hus((size_t)k * rank, 0.0);
        to_host(hu.data(), o.u.get(), hu.size(), s);
        for (int r = 0; r < k; ++r)
            for (int c = 0; c < rank; ++c) hus[(size_t)r * rank + c] = hu[(size_t)r * k + cols[c]];
        DevBuf<double> us((size_t)k * rank);
        to_device(us.get(), hus.data(), hus.size(), s);
        block_mul_f64(n, y, kp, k, us, rank, nullptr, 0, 0, nullptr, evecs_dev, rank, s);
    } else {
        LargeOrth o(n, k, kp);
        for (int it = 0; it < nbiter; ++it) {
            apply_m(om, y);
            o.orth(n, y, kp, s);
            apply_m(y, om);
            o.orth(n, om, kp, s);
        }
        apply_m(om, y);
        o.orth(n, y, kp, s);       // Q
        apply_m(y, om);            // M Q
        o.gram(n, y, om, kp, s);   // B = Q^T M Q
        o.eig(s);                  // ascending; U_B in the columns of o.g
        to_host(hw.data(), o.w.get(), k, s);
        // the selected columns of U_B, in output order (column-major k x rank)
        DevBuf<double> us((size_t)k * rank);
        for (int c = 0; c < rank; ++c)
            SI_HIP(hipMemcpyAsync(us.get() + (size_t)c * k, o.g.get() + (size_t)cols[c] * k,
                                  sizeof(double) * k, hipMemcpyDeviceToDevice, s));
        o.mul(n, y, kp, us, k, rank, evecs_dev, rank, s);
    }
    for (int c = 0; c < rank; ++c) evals_host[c] = 2.0 - hw[cols[c]];
    SI_HIP(hipStreamSynchronize(s));
    st.iterations = nbiter;
}

}  // namespace si
