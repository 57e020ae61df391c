// residual_noalloc.hip -- si_max_eigenpair_residual_noalloc: the FP64 residual
// max_j ||L v_j - lambda_j v_j|| / max(||v_j||, 1e-300) (the semantics of
// src/metrics.rs:57-92, as si_max_eigenpair_residual computes it) for callers
// that cannot allocate device memory.
//
// The call allocates nothing.  Its only device memory beyond the caller's
// inputs and the Laplacian is `scratch`, a file-scope array of
// kNaScratchDoubles doubles (1 MiB + 1 KiB) that every GPU the library runs on
// holds until the process ends.  Its device address is looked up once per
// device, with that device current, when a context is created on it
// (hipGetSymbolAddress); host code uses only that address, never the array's
// name.  Every call on one GPU shares that GPU's array: calls that overlap on
// one GPU (two streams, two threads) race.  Nothing in the array needs to
// start at zero: each call writes every partial sum it later reads.
//
// Two kernels:
//   k_na_partial   one thread per row (grid-stride, at most kNaGroups
//                  workgroups), up to 8 columns per launch: (L v)_i gathered
//                  straight from the caller's block through the locality
//                  permutation, d = (L v)_i - lambda v_i, per-thread sums of
//                  d^2 and v_i^2, a fixed-order workgroup reduction into
//                  scratch[group][column];
//   k_na_final     one workgroup: per-column sums over the groups in order,
//                  the residual with the 1e-300 norm floor, NaN-propagating
//                  maximum over the columns.
// Sums have a fixed order, so the result is deterministic.
#include <mutex>

#include "si_abi.hpp"

namespace {

constexpr int kNaGroups = 1024;   // workgroups of the partial-sum pass
constexpr int kNaMaxK = 64;       // columns per call
constexpr int kNaChunk = 8;       // columns per partial-sum launch
constexpr int kNaThreads = 256;
// [kNaGroups][2 * kNaMaxK] partial sums (d^2 in the first kNaMaxK slots, v^2 in
// the next), then kNaMaxK column residuals and the maximum
constexpr size_t kNaPartDoubles = (size_t)kNaGroups * 2 * kNaMaxK;
constexpr size_t kNaScratchDoubles = kNaPartDoubles + kNaMaxK + 64;

__device__ __align__(256) double scratch[kNaScratchDoubles];

constexpr int kMaxDevices = 64;
double* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

struct Lambdas {
    double v[kNaMaxK];
};

__global__ __launch_bounds__(kNaThreads) void k_na_partial(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
    const double* __restrict__ l64, const int32_t* __restrict__ perm,
    const double* __restrict__ v, int k, int c0, int cw, int64_t n, Lambdas lam,
    double* __restrict__ part) {
    double sdd[kNaChunk], svv[kNaChunk];
#pragma unroll
    for (int c = 0; c < kNaChunk; ++c) sdd[c] = svv[c] = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kNaThreads;
    for (int64_t i = (int64_t)blockIdx.x * kNaThreads + threadIdx.x; i < n; i += stride) {
        const int64_t vi = perm ? perm[i] : i;
        const double* xr = v + vi * k + c0;
        double xi[kNaChunk], acc[kNaChunk];
#pragma unroll
        for (int c = 0; c < kNaChunk; ++c) {
            xi[c] = c < cw ? xr[c] : 0.0;
            acc[c] = 0.0;
        }
        // l64 holds L's off-diagonal entries; the unit diagonal is added last,
        // as the FP64 SpMM's epilogue does
        for (int32_t p = rp[i]; p < rp[i + 1]; ++p) {
            const int32_t j = col[p];
            const int64_t vj = perm ? perm[j] : j;
            const double w = l64[p];
            const double* xj = v + vj * k + c0;
#pragma unroll
            for (int c = 0; c < kNaChunk; ++c)
                if (c < cw) acc[c] += w * xj[c];
        }
#pragma unroll
        for (int c = 0; c < kNaChunk; ++c) {
            if (c < cw) {
                const double d = (xi[c] + acc[c]) - lam.v[c0 + c] * xi[c];
                sdd[c] += d * d;
                svv[c] += xi[c] * xi[c];
            }
        }
    }
    // fixed-order reduction: wave butterflies, then the 4 wave sums in order
    __shared__ double red[kNaThreads / 64][2 * kNaChunk];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int c = 0; c < kNaChunk; ++c) {
            sdd[c] += __shfl_xor(sdd[c], off, 64);
            svv[c] += __shfl_xor(svv[c], off, 64);
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < kNaChunk; ++c) {
            red[wave][c] = sdd[c];
            red[wave][kNaChunk + c] = svv[c];
        }
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < 2 * cw) {
        const int slot = t < cw ? t : kNaChunk + (t - cw);
        const double s = ((red[0][slot] + red[1][slot]) + red[2][slot]) + red[3][slot];
        const int c = t < cw ? t : t - cw;
        const int off = t < cw ? 0 : kNaMaxK;
        part[(size_t)blockIdx.x * 2 * kNaMaxK + off + c0 + c] = s;
    }
}

__global__ __launch_bounds__(kNaMaxK) void k_na_final(const double* __restrict__ part, int groups,
                                                      int k, double* __restrict__ result) {
    __shared__ double res[kNaMaxK];
    const int c = threadIdx.x;
    if (c < k) {
        double dd = 0.0, vv = 0.0;
        for (int g = 0; g < groups; ++g) {
            dd += part[(size_t)g * 2 * kNaMaxK + c];
            vv += part[(size_t)g * 2 * kNaMaxK + kNaMaxK + c];
        }
        double r;
        if (dd != dd || vv != vv) {
            r = __builtin_nan("");
        } else {
            double vn = sqrt(vv);
            if (vn < 1e-300) vn = 1e-300;
            r = sqrt(dd) / vn;
        }
        res[c] = r;
        result[c] = r;
    }
    __syncthreads();
    if (c == 0) {
        double q = 0.0;
        for (int j = 0; j < k; ++j) {
            const double r = res[j];
            q = (q != q || r != r) ? __builtin_nan("") : (r > q ? r : q);
        }
        result[kNaMaxK] = q;
    }
}

}  // namespace

namespace si {

// Device address of `scratch` on the current device (hipGetSymbolAddress), kept
// per device.  Called by si_context_create with the context's device current.
void residual_noalloc_setup(int device) {
    SI_CHECK(device >= 0 && device < kMaxDevices, kErrUnsupported,
             "residual_noalloc: device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    SI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SI_CHECK(p != nullptr, kErrHip, "residual_noalloc: no device address for the scratch array");
    g_scratch_ptr[device] = static_cast<double*>(p);
}

}  // namespace si

extern "C" si_status si_max_eigenpair_residual_noalloc(si_context* ctx, const si_laplacian* lap,
                                                       const double* evals, const double* evecs,
                                                       int32_t k, double* out) {
    if (!ctx || !lap || !evals || !evecs || !out) return SI_ERR_INVALID_ARGUMENT;
    try {
        const int dev = ctx->c.device;
        SI_HIP(hipSetDevice(dev));
        SI_CHECK(k >= 1 && k <= kNaMaxK, si::kErrInvalidArgument,
                 "max_eigenpair_residual_noalloc: k = " + std::to_string(k) +
                     " is outside [1, 64], the columns its static scratch holds");
        SI_CHECK(lap->device == dev, si::kErrInvalidArgument,
                 "max_eigenpair_residual_noalloc: Laplacian and context are on different devices");
        double* sc = nullptr;
        {
            std::lock_guard<std::mutex> lock(g_scratch_mu);
            if (dev >= 0 && dev < kMaxDevices) sc = g_scratch_ptr[dev];
        }
        SI_CHECK(sc != nullptr, si::kErrInternal,
                 "max_eigenpair_residual_noalloc: scratch not set up on this device");
        const si::DevLaplacian& L = lap->L;
        hipStream_t s = ctx->c.stream;
        Lambdas lam{};
        for (int c = 0; c < k; ++c) lam.v[c] = evals[c];
        int64_t groups = si::ceil_div(L.n, (int64_t)kNaThreads);
        if (groups > kNaGroups) groups = kNaGroups;
        if (groups < 1) groups = 1;
        double* part = sc;
        double* result = sc + kNaPartDoubles;
        for (int c0 = 0; c0 < k; c0 += kNaChunk) {
            const int cw = (k - c0) < kNaChunk ? (k - c0) : kNaChunk;
            hipLaunchKernelGGL(k_na_partial, dim3((unsigned)groups), dim3(kNaThreads), 0, s,
                               (const int32_t*)L.rowptr, (const int32_t*)L.col,
                               (const double*)L.l64, (const int32_t*)L.perm, evecs, (int)k, c0,
                               cw, L.n, lam, part);
            SI_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_na_final, dim3(1), dim3(kNaMaxK), 0, s, (const double*)part,
                           (int)groups, (int)k, result);
        SI_LAUNCH_CHECK();
        SI_HIP(hipMemcpyAsync(out, result + kNaMaxK, sizeof(double), hipMemcpyDeviceToHost, s));
        SI_HIP(hipStreamSynchronize(s));
        ctx->c.last_error.clear();
        return SI_OK;
    } catch (const si::Error& e) {
        ctx->c.last_error = e.what();
        return (si_status)e.code;
    } catch (const std::exception& e) {
        ctx->c.last_error = e.what();
        return SI_ERR_INTERNAL;
    }
}
