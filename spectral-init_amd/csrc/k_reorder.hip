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

// Rows of up to kRowSortMax entries are ranked in registers, four entries per
// lane; longer rows go to a segmented sort.
constexpr int kRowSortMax = 256;

// new row i = original row perm[i], columns renumbered through iperm and
// sorted (stable: equal columns keep their order); one wavefront per row,
// grid-strided.  Entry pos of a row sits in lane pos % 64, register pos / 64;
// its rank is the number of entries before it in the sorted row.  A row above
// kRowSortMax entries is copied unsorted and listed in longs[1 ..] (count in
// longs[0]).
__global__ __launch_bounds__(256) void k_perm_rows(int64_t n, const int64_t* __restrict__ rp,
                                                   const int32_t* __restrict__ col,
                                                   const float* __restrict__ w,
                                                   const int32_t* __restrict__ perm,
                                                   const int32_t* __restrict__ iperm,
                                                   const int32_t* __restrict__ rp2,
                                                   int32_t* __restrict__ c2, float* __restrict__ w2,
                                                   int32_t* __restrict__ longs) {
    constexpr int KPL = kRowSortMax / 64;
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n; i += waves) {
        const int32_t v = perm[i];
        const int64_t b = rp[v];
        const int32_t len = (int32_t)(rp[v + 1] - b);
        const int64_t o = rp2[i];
        if (len <= 64) {
            int32_t key = INT32_MAX;
            float val = 0.0f;
            if (lane < len) {
                key = iperm[col[b + lane]];
                val = w[b + lane];
            }
            int32_t r = 0;
            for (int t = 0; t < len; ++t) {
                const int32_t kt = __shfl(key, t);
                r += (kt < key || (kt == key && t < lane)) ? 1 : 0;
            }
            if (lane < len) {
                c2[o + r] = key;
                w2[o + r] = val;
            }
        } else if (len <= kRowSortMax) {
            int32_t key[KPL], r[KPL];
            float val[KPL];
#pragma unroll
            for (int q = 0; q < KPL; ++q) {
                const int32_t pos = lane + 64 * q;
                key[q] = INT32_MAX;
                val[q] = 0.0f;
                r[q] = 0;
                if (pos < len) {
                    key[q] = iperm[col[b + pos]];
                    val[q] = w[b + pos];
                }
            }
            for (int t = 0; t < len; ++t) {
                const int qt = t >> 6;
                int32_t src = key[0];
#pragma unroll
                for (int q = 1; q < KPL; ++q)
                    if (qt == q) src = key[q];
                const int32_t kt = __shfl(src, t & 63);
#pragma unroll
                for (int q = 0; q < KPL; ++q)
                    r[q] += (kt < key[q] || (kt == key[q] && t < lane + 64 * q)) ? 1 : 0;
            }
#pragma unroll
            for (int q = 0; q < KPL; ++q)
                if (lane + 64 * q < len) {
                    c2[o + r[q]] = key[q];
                    w2[o + r[q]] = val[q];
                }
        } else {
            for (int32_t p = lane; p < len; p += 64) {
                c2[o + p] = iperm[col[b + p]];
                w2[o + p] = w[b + p];
            }
            if (lane == 0) longs[1 + atomicAdd(&longs[0], 1)] = (int32_t)i;
        }
    }
}

// bounds of the listed long rows in the new CSR, and their lengths (len[count]
// = 0, the exclusive-scan pad)
__global__ void k_long_bounds(int32_t count, const int32_t* __restrict__ longs,
                              const int32_t* __restrict__ rp2, int32_t* __restrict__ beg,
                              int32_t* __restrict__ len) {
    const int32_t j = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j > count) return;
    if (j == count) {
        len[count] = 0;
        return;
    }
    const int32_t i = longs[1 + j];
    beg[j] = rp2[i];
    len[j] = rp2[i + 1] - rp2[i];
}

// long row j between the new CSR (at beg[j]) and a compact buffer (at off[j]);
// one wavefront per row
template <bool TO_COMPACT>
__global__ __launch_bounds__(256) void k_long_move(int32_t count, const int32_t* __restrict__ beg,
                                                   const int32_t* __restrict__ off,
                                                   int32_t* __restrict__ c2, float* __restrict__ w2,
                                                   int32_t* __restrict__ ck, float* __restrict__ wk) {
    const int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (j >= count) return;
    const int32_t b = beg[j], o = off[j], len = off[j + 1] - off[j];
    for (int32_t t = (int32_t)(threadIdx.x & 63); t < len; t += 64) {
        if (TO_COMPACT) {
            ck[o + t] = c2[b + t];
            wk[o + t] = w2[b + t];
        } else {
            c2[b + t] = ck[o + t];
            w2[b + t] = wk[o + t];
        }
    }
}

}  // namespace

void permute_csr_device(int64_t n, int64_t nnz, const int64_t* rp, const int32_t* col,
                        const float* w, const int32_t* perm, const int32_t* iperm, int32_t* rp2,
                        int32_t* c2, float* w2, hipStream_t s) {
    SI_CHECK(n >= 1 && nnz >= 0 && nnz < (int64_t)INT32_MAX, kErrUnsupported,
             "permute_csr: sizes out of range");
    StreamBuf<int32_t> lens(n + 1, s);
    hipLaunchKernelGGL(k_perm_lens, dim3((unsigned)ceil_div(n + 1, 256)), dim3(256), 0, s, n, rp,
                       perm, lens.get());
    SI_LAUNCH_CHECK();
    size_t tb = 0;
    SI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, lens.get(), rp2, (int)(n + 1), s));
    StreamBuf<char> tmp(tb, s);
    SI_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.get(), tb, lens.get(), rp2, (int)(n + 1), s));
    if (nnz == 0) {
        SI_HIP(hipStreamSynchronize(s));
        return;
    }
    StreamBuf<int32_t> longs(n + 1, s);
    SI_HIP(hipMemsetAsync(longs.get(), 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_perm_rows, dim3((unsigned)std::min<int64_t>(ceil_div(n, 4), 4096)),
                       dim3(256), 0, s, n, rp, col, w, perm, iperm, rp2, c2, w2, longs.get());
    SI_LAUNCH_CHECK();
    int32_t count = 0;
    SI_HIP(hipMemcpyAsync(&count, longs.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SI_HIP(hipStreamSynchronize(s));
    SI_CHECK(count >= 0 && count <= n, kErrInternal, "permute_csr: bad long-row count");
    if (count > 0) {
        // rows above kRowSortMax entries: copied into a compact buffer, sorted there by
        // a segmented sort, copied back
        StreamBuf<int32_t> beg(count, s), len(count + 1, s), off(count + 1, s);
        hipLaunchKernelGGL(k_long_bounds, dim3((unsigned)ceil_div(count + 1, 256)), dim3(256), 0, s,
                           count, longs.get(), rp2, beg.get(), len.get());
        SI_LAUNCH_CHECK();
        size_t tb1 = 0;
        SI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb1, len.get(), off.get(), count + 1, s));
        StreamBuf<char> tmp1(tb1, s);
        SI_HIP(hipcub::DeviceScan::ExclusiveSum(tmp1.get(), tb1, len.get(), off.get(), count + 1, s));
        int32_t total = 0;
        SI_HIP(hipMemcpyAsync(&total, off.get() + count, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        SI_HIP(hipStreamSynchronize(s));
        SI_CHECK(total > 0 && total <= nnz, kErrInternal, "permute_csr: bad long-row total");
        StreamBuf<int32_t> ka(total, s), kb(total, s);
        StreamBuf<float> va(total, s), vb(total, s);
        const dim3 rows_grid((unsigned)ceil_div((int64_t)count * 64, 256));
        hipLaunchKernelGGL(k_long_move<true>, rows_grid, dim3(256), 0, s, count, beg.get(),
                           off.get(), c2, w2, ka.get(), va.get());
        SI_LAUNCH_CHECK();
        size_t tb2 = 0;
        SI_HIP(hipcub::DeviceSegmentedSort::StableSortPairs(nullptr, tb2, ka.get(), kb.get(), va.get(),
                                                            vb.get(), total, count, off.get(),
                                                            off.get() + 1, s));
        StreamBuf<char> tmp2(tb2, s);
        SI_HIP(hipcub::DeviceSegmentedSort::StableSortPairs(tmp2.get(), tb2, ka.get(), kb.get(),
                                                            va.get(), vb.get(), total, count,
                                                            off.get(), off.get() + 1, s));
        hipLaunchKernelGGL(k_long_move<false>, rows_grid, dim3(256), 0, s, count, beg.get(),
                           off.get(), c2, w2, kb.get(), vb.get());
        SI_LAUNCH_CHECK();
    }
    SI_HIP(hipStreamSynchronize(s));
}

}  // namespace si
