// solver_sinv.cpp -- lobpcg_sinv_solve (src/solvers/sinv.rs:106-189) on the
// device: escalation level 2, shift-and-invert LOBPCG.
//
// M = L + SINV_SHIFT I is factorised once (banded Cholesky in the device's node
// order, k_band.hip; the reference uses a sparse Cholesky, sinv.rs:125), then
// LOBPCG runs on M^-1 for its k largest eigenvalues mu (Order::Largest,
// sinv.rs:146-153).  lobpcg_core finds the smallest eigenvalues of its
// operator, so the operator here is -M^-1 and mu = -nu.  Start block, tolerance,
// iteration cap and acceptance follow sinv.rs:128-167; eigenvalues are
// recovered as lambda = 1/mu - SINV_SHIFT (sinv.rs:170-171).  Ascending nu is
// descending mu, i.e. ascending lambda, so the columns come out in the
// reference's sorted order (sinv.rs:173-186).
#include <cmath>

#include "si_host.hpp"

namespace si {

bool lobpcg_sinv_solve_device(const DevLaplacian& L, int n_components, uint64_t seed,
                              double* evals_host, double* evecs_dev, SolveStats& st,
                              hipStream_t s) {
    const int64_t n = L.n;
    const int k = n_components + 1;
    st = SolveStats{};
    st.block_size = k;
    if (n_components <= 0 || k >= n) return false;  // sinv.rs:118-120
    BandFactor F;
    if (!band_factor(L, kSinvShift, F, s)) return false;  // sinv.rs:125
    const size_t nk = (size_t)n * k;
    DevBuf<double> x(nk), g(1), ws(gram_workspace_doubles(n, 1, 1));
    // N(0,1) start with the normalised sqrt_deg in column 0 (sinv.rs:128-134)
    fill_normal_f64(x, n, k, k, seed, kStreamLobpcgInit, s, L.perm);
    gram_f64(n, L.sqrt_deg, 1, 1, L.sqrt_deg, 1, 1, g, ws, s);
    double sn2 = 0.0;
    to_host(&sn2, g.get(), 1, s);
    const double sn = std::sqrt(sn2);
    if (sn > 0.0) set_column_f64(x, k, n, 0, L.sqrt_deg, 1.0 / sn, s);

    const BlockOperator op = [&](const double* in, double* out, int kk, int ld) {
        band_solve(F, in, ld, out, ld, kk, -1.0, s);
    };
    const double tol = (double)(float)1e-8;  // SINV_LINFA_TOL as f32 (sinv.rs:150)
    const int maxiter = (int)std::min<int64_t>(5 * n, 300);
    std::vector<double> nu(k), rn(k);
    int64_t calls = 0;
    const bool conv = lobpcg_core(n, op, x, k, tol, maxiter, nu.data(), rn.data(), st.iterations,
                                  calls, s, /*robust=*/true);
    bool accept = conv;
    if (!conv) {  // sinv.rs:160-164: partial results with every rnorm below the threshold
        accept = true;
        for (int c = 0; c < k; ++c)
            if (!(rn[c] < kSinvQualityThreshold)) accept = false;
    }
    if (!accept) return false;
    for (int c = 0; c < k; ++c) evals_host[c] = 1.0 / (-nu[c]) - kSinvShift;
    SI_HIP(hipMemcpyAsync(evecs_dev, x.get(), sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
    SI_HIP(hipStreamSynchronize(s));
    st.converged = conv;
    return true;
}

}  // namespace si
