// k_reorder.hip -- the locality order applied on the device: the permuted CSR
// P W P^T of the graph (rows in the new order, columns renumbered and sorted
// within each row), from the original CSR already resident on the device.
#include <hipcub/hipcub.hpp>

#include "si_host.hpp"

namespace si {

namespace {

// lens[i] = length of original row perm[i]; lens[n] = 0 (exclusive-scan pad)
__global__ void k_perm_lens(int64_t n, const int64_t* __restrict__ rp,
                            const int32_t* __restrict__ perm, int32_t* __restrict__ lens) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { lens[n] = 0; return; }
    const int32_t v = perm[i];
    lens[i] = (int32_t)(rp[v + 1] - rp[v]);
}

// new row i = original row perm[i], columns renumbered through iperm (order
// kept; sorted afterwards); 16 lanes per row, 16 rows per workgroup at a time,
// grid-strided
__global__ __launch_bounds__(256) void k_perm_fill(int64_t n, const int64_t* __restrict__ rp,
                                                   const int32_t* __restrict__ col,
                                                   const float* __restrict__ w,
                                                   const int32_t* __restrict__ perm,
                                                   const int32_t* __restrict__ iperm,
                                                   const int32_t* __restrict__ rp2,
                                                   int32_t* __restrict__ c2, float* __restrict__ w2) {
    const int sub = threadIdx.x & 15;
    for (int64_t base = (int64_t)blockIdx.x * 16; base < n; base += (int64_t)gridDim.x * 16) {
        const int64_t i = base + (threadIdx.x >> 4);
        if (i >= n) continue;
        const int32_t v = perm[i];
        const int64_t b = rp[v], e = rp[v + 1];
        const int64_t o = rp2[i] - b;
        for (int64_t p = b + sub; p < e; p += 16) {
            c2[o + p] = iperm[col[p]];
            w2[o + p] = w[p];
        }
    }
}

}  // namespace

void permute_csr_device(int64_t n, int64_t nnz, const int64_t* rp, const int32_t* col,
                        const float* w, const int32_t* perm, const int32_t* iperm, int32_t* rp2,
                        int32_t* c2, float* w2, hipStream_t s) {
    SI_CHECK(n >= 1 && nnz >= 0 && nnz < (int64_t)INT32_MAX, kErrUnsupported,
             "permute_csr: sizes out of range");
    DevBuf<int32_t> lens(n + 1);
    hipLaunchKernelGGL(k_perm_lens, dim3((unsigned)ceil_div(n + 1, 256)), dim3(256), 0, s, n, rp,
                       perm, lens.get());
    SI_LAUNCH_CHECK();
    size_t tb = 0;
    SI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, lens.get(), rp2, (int)(n + 1), s));
    DevBuf<char> tmp(tb);
    SI_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.get(), tb, lens.get(), rp2, (int)(n + 1), s));
    if (nnz == 0) {
        SI_HIP(hipStreamSynchronize(s));
        return;
    }
    DevBuf<int32_t> ck(nnz);
    DevBuf<float> wk(nnz);
    hipLaunchKernelGGL(k_perm_fill, dim3((unsigned)std::min<int64_t>(ceil_div(n, 16), 4096)),
                       dim3(256), 0, s, n, rp,
                       col, w, perm, iperm, rp2, ck.get(), wk.get());
    SI_LAUNCH_CHECK();
    // sort every row by its new column numbers (segments = the new rows)
    size_t tb2 = 0;
    SI_HIP(hipcub::DeviceSegmentedSort::StableSortPairs(nullptr, tb2, ck.get(), c2, wk.get(), w2,
                                                        (int)nnz, (int)n, rp2, rp2 + 1, s));
    DevBuf<char> tmp2(tb2);
    SI_HIP(hipcub::DeviceSegmentedSort::StableSortPairs(tmp2.get(), tb2, ck.get(), c2, wk.get(), w2,
                                                        (int)nnz, (int)n, rp2, rp2 + 1, s));
    SI_HIP(hipStreamSynchronize(s));
}

}  // namespace si
