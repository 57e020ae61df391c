// k_rcm.hip -- connected components and the reverse Cuthill-McKee locality
// order of a symmetric CSR pattern, on the device.
//
// Components (the labelling of find_components, src/components.rs:7-44, for
// symmetric graphs): min-label hooking in rounds.  Every node keeps a pointer
// f[v] <= v inside its component; a round hooks f[f[u]] and f[u] to the
// grandparent f[f[v]] of every neighbour v (device-scope atomic minima) and
// then halves the pointer paths.  Pointers only decrease, so a round without
// any change sees the values of the previous kernel boundary: at that fixed
// point f is constant on every component and equal to its smallest node.
// Components are numbered by their smallest node -- the order in which a BFS
// sweep over the nodes in id order meets them.
//
// RCM (the order of rcm_order in reorder.cpp, bit for bit): for every
// component two BFS sweeps find a pseudo-peripheral start (the first
// minimum-degree node, in queue order, of the last level) and a third sweep
// visits each parent's new neighbours by increasing (degree, id); the final
// order is reversed.  All components sweep together, level-synchronously, in
// one global queue: level L of every component (in component order), then
// level L + 1.  Within a component that is exactly its own queue order, so a
// node of the next level belongs to the first adjacent node of the level
// (minimum queue position), found as a 64-bit atomic minimum of (position,
// entry) over the level's rows -- the entry part keeps only the first of
// repeated columns.  Per level: claim, count (one wavefront per level node),
// exclusive scan, emit (ballot ranks in adjacency order; in the last sweep
// groups of <= 64 new nodes are ranked by (degree, id) in registers, larger
// groups by a segmented radix sort).  Each level costs one 8-byte readback.
// A stable radix sort by component label finally turns the global queue into
// the concatenation of the per-component orders.
#include <hipcub/hipcub.hpp>

#include <climits>
#include <vector>

#include "si_host.hpp"

namespace si {

namespace {

constexpr uint64_t kNoClaim = ~0ull;

// ----------------------------------------------------------------- components

__global__ void k_cc_init(int64_t n, int32_t* __restrict__ f) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = (int32_t)i;
}

// hooking: f[f[u]] and f[u] to min(f[f[v]]) over the neighbours v of u
__global__ void k_cc_hook(int64_t n, const int64_t* __restrict__ rp, const int32_t* __restrict__ ix,
                          int32_t* __restrict__ f, int* __restrict__ changed) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const int64_t b = rp[u], e = rp[u + 1];
    if (b == e) return;
    const int32_t fu = f[u];
    int32_t best = INT32_MAX;
    for (int64_t p = b; p < e; ++p) best = min(best, f[f[ix[p]]]);
    bool ch = false;
    if (best < fu) {
        ch |= atomicMin(&f[fu], best) > best;
        ch |= atomicMin(&f[u], best) > best;
    }
    if (ch) *changed = 1;
}

// path halving: f[u] = min(f[u], f[f[u]])  (thread u is the only writer of f[u])
__global__ void k_cc_shortcut(int64_t n, int32_t* __restrict__ f, int* __restrict__ changed) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const int32_t fu = f[u];
    const int32_t g = f[fu];
    if (g < fu) {
        f[u] = g;
        *changed = 1;
    }
}

__global__ void k_cc_roots(int64_t n, const int32_t* __restrict__ f, int32_t* __restrict__ isroot) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) isroot[i] = f[i] == (int32_t)i ? 1 : 0;
    else if (i == n) isroot[n] = 0;
}

__global__ void k_cc_label(int64_t n, const int32_t* __restrict__ f, const int32_t* __restrict__ rank,
                           int32_t* __restrict__ lab, int32_t* __restrict__ root) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t c = rank[f[i]];
    lab[i] = c;
    if (root && f[i] == (int32_t)i) root[c] = (int32_t)i;
}

// ----------------------------------------------------------------- RCM sweeps

__device__ __forceinline__ uint64_t claim_key(int64_t q, int64_t p) {
    return ((uint64_t)q << 32) | (uint64_t)(uint32_t)p;
}

// level 0 of a sweep: queue[c] = start[c]
__global__ void k_rcm_start(int32_t nc, const int32_t* __restrict__ start, int32_t* __restrict__ queue,
                            int32_t* __restrict__ mark, int32_t* __restrict__ lvl, int32_t epoch) {
    const int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= nc) return;
    const int32_t v = start[c];
    queue[c] = v;
    mark[v] = epoch;
    lvl[c] = 0;
}

// claim: key[u] = min (position, entry) over the level's entries (q, p) with
// unvisited column u; one wavefront per level node
__global__ __launch_bounds__(256) void k_rcm_claim(int64_t ls, int64_t m, const int64_t* __restrict__ rp,
                                                   const int32_t* __restrict__ ix,
                                                   const int32_t* __restrict__ queue,
                                                   const int32_t* __restrict__ mark, int32_t epoch,
                                                   unsigned long long* __restrict__ key) {
    const int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (j >= m) return;
    const int64_t q = ls + j;
    const int32_t v = queue[q];
    for (int64_t p = rp[v] + lane; p < rp[v + 1]; p += 64) {
        const int32_t u = ix[p];
        if (mark[u] != epoch) atomicMin(&key[u], (unsigned long long)claim_key(q, p));
    }
}

// count: cnt[j] = entries of level node j that own their column
__global__ __launch_bounds__(256) void k_rcm_count(int64_t ls, int64_t m, const int64_t* __restrict__ rp,
                                                   const int32_t* __restrict__ ix,
                                                   const int32_t* __restrict__ queue,
                                                   const int32_t* __restrict__ mark, int32_t epoch,
                                                   const unsigned long long* __restrict__ key,
                                                   int32_t* __restrict__ cnt) {
    const int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (j > m) return;
    if (j == m) {
        if (lane == 0) cnt[m] = 0;
        return;
    }
    const int64_t q = ls + j;
    const int32_t v = queue[q];
    int c = 0;
    for (int64_t p = rp[v] + lane; p < rp[v + 1]; p += 64) {
        const int32_t u = ix[p];
        c += (mark[u] != epoch && key[u] == claim_key(q, p)) ? 1 : 0;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane == 0) cnt[j] = c;
}

// emit: the owned columns of level node j go to queue[le + off[j] + rank], in
// adjacency order; in the sorted sweep a group of <= 64 is ranked by
// (degree, id) here and every group's keys also go to skey (for the
// segmented sort of a level with a larger group, flagged in *big)
__global__ __launch_bounds__(256) void k_rcm_emit(int64_t ls, int64_t le, int64_t m,
                                                  const int64_t* __restrict__ rp,
                                                  const int32_t* __restrict__ ix,
                                                  int32_t* __restrict__ queue,
                                                  int32_t* __restrict__ mark, int32_t epoch,
                                                  unsigned long long* __restrict__ key,
                                                  const int32_t* __restrict__ off,
                                                  int32_t* __restrict__ lvl, int32_t level,
                                                  const int32_t* __restrict__ lab,
                                                  int32_t* __restrict__ maxlev, int sorted,
                                                  unsigned long long* __restrict__ skey,
                                                  int* __restrict__ big) {
    __shared__ unsigned long long grp[4][64];
    const int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    if (j >= m) return;
    const int64_t q = ls + j;
    const int32_t v = queue[q];
    const int32_t o = off[j];
    const int32_t g = off[j + 1] - o;
    if (g == 0) return;
    int32_t run = 0;
    for (int64_t p0 = rp[v]; p0 < rp[v + 1]; p0 += 64) {
        const int64_t p = p0 + lane;
        int32_t u = -1;
        bool own = false;
        if (p < rp[v + 1]) {
            u = ix[p];
            own = mark[u] != epoch && key[u] == claim_key(q, p);
        }
        const uint64_t bal = __ballot(own);
        if (own) {
            const int32_t r = run + __popcll(bal & ((1ull << lane) - 1ull));
            const uint64_t dk = ((uint64_t)(rp[u + 1] - rp[u]) << 32) | (uint32_t)u;
            if (!sorted) {
                queue[le + o + r] = u;
            } else if (g <= 64) {
                grp[w][r] = dk;
            } else {
                skey[o + r] = dk;
            }
            lvl[le + o + r] = level;
            mark[u] = epoch;
            key[u] = kNoClaim;
        }
        run += __popcll(bal);
    }
    if (lane == 0) atomicMax(&maxlev[lab[v]], level);
    if (sorted) {
        if (g <= 64) {
            __syncwarp();
            if (lane < g) {
                const unsigned long long mine = grp[w][lane];
                int32_t r = 0;
                for (int t = 0; t < g; ++t) r += grp[w][t] < mine ? 1 : 0;
                queue[le + o + r] = (int32_t)(uint32_t)mine;
                skey[o + r] = mine;
            }
        } else if (lane == 0) {
            *big = 1;
        }
    }
}

// after the segmented sort of a level: queue[le + i] = id of sorted key i
__global__ void k_rcm_scatter(int64_t le, int64_t cnt, const unsigned long long* __restrict__ sk,
                              int32_t* __restrict__ queue) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) queue[le + i] = (int32_t)(uint32_t)sk[i];
}

// first minimum-degree node (queue order) of each component's last level
__global__ void k_rcm_last_level(int64_t n, const int64_t* __restrict__ rp,
                                 const int32_t* __restrict__ queue, const int32_t* __restrict__ lvl,
                                 const int32_t* __restrict__ lab, const int32_t* __restrict__ maxlev,
                                 unsigned long long* __restrict__ best) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int32_t v = queue[q];
    const int32_t c = lab[v];
    if (lvl[q] != maxlev[c]) return;
    atomicMin(&best[c], (unsigned long long)(((uint64_t)(rp[v + 1] - rp[v]) << 32) | (uint64_t)q));
}

__global__ void k_rcm_next_start(int32_t nc, const unsigned long long* __restrict__ best,
                                 const int32_t* __restrict__ queue, int32_t* __restrict__ start) {
    const int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c < nc) start[c] = queue[(uint32_t)best[c]];
}

__global__ void k_fill_u64(unsigned long long* __restrict__ a, int64_t n, unsigned long long v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_fill_i32(int32_t* __restrict__ a, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_queue_labels(int64_t n, const int32_t* __restrict__ queue,
                               const int32_t* __restrict__ lab, int32_t* __restrict__ ql) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ql[i] = lab[queue[i]];
}

__global__ void k_rcm_reverse(int64_t n, const int32_t* __restrict__ order, int32_t* __restrict__ perm,
                              int32_t* __restrict__ iperm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t v = order[n - 1 - i];
    perm[i] = v;
    iperm[v] = (int32_t)i;
}

inline dim3 grid_of(int64_t threads) { return dim3((unsigned)ceil_div(threads, 256)); }

}  // namespace

int64_t components_device(int64_t n, const int64_t* rp, const int32_t* ix, int32_t* lab,
                          int32_t* root, hipStream_t s) {
    SI_CHECK(n >= 1 && n < (int64_t)INT32_MAX, kErrUnsupported, "components: n out of range");
    DevBuf<int32_t> f(n), rk(n + 1);
    DevBuf<int> flag(1);
    hipLaunchKernelGGL(k_cc_init, grid_of(n), dim3(256), 0, s, n, f.get());
    SI_LAUNCH_CHECK();
    // every changing round lowers some pointer, so n rounds bound the loop; in
    // practice it ends after O(log diameter) rounds
    int64_t rounds = 0;
    for (;;) {
        SI_CHECK(rounds++ <= n + 1, kErrInternal, "components: no fixed point");
        SI_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), s));
        hipLaunchKernelGGL(k_cc_hook, grid_of(n), dim3(256), 0, s, n, rp, ix, f.get(), flag.get());
        hipLaunchKernelGGL(k_cc_shortcut, grid_of(n), dim3(256), 0, s, n, f.get(), flag.get());
        SI_LAUNCH_CHECK();
        int ch = 0;
        SI_HIP(hipMemcpyAsync(&ch, flag.get(), sizeof(int), hipMemcpyDeviceToHost, s));
        SI_HIP(hipStreamSynchronize(s));
        if (!ch) break;
    }
    DevBuf<int32_t> isroot(n + 1);
    hipLaunchKernelGGL(k_cc_roots, grid_of(n + 1), dim3(256), 0, s, n, f.get(), isroot.get());
    size_t tb = 0;
    SI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, isroot.get(), rk.get(), (int)(n + 1), s));
    DevBuf<char> tmp(tb);
    SI_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.get(), tb, isroot.get(), rk.get(), (int)(n + 1), s));
    hipLaunchKernelGGL(k_cc_label, grid_of(n), dim3(256), 0, s, n, f.get(), rk.get(), lab, root);
    SI_LAUNCH_CHECK();
    int32_t nc = 0;
    SI_HIP(hipMemcpyAsync(&nc, rk.get() + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SI_HIP(hipStreamSynchronize(s));
    return nc;
}

int64_t rcm_order_device(int64_t n, const int64_t* rp, const int32_t* ix, int32_t* perm,
                         int32_t* iperm, hipStream_t s, int32_t* lab_out) {
    SI_CHECK(n >= 1 && n < (int64_t)INT32_MAX, kErrUnsupported, "rcm: n out of range");
    DevBuf<int32_t> lab_own(lab_out ? 0 : n), root(n);
    int32_t* lab = lab_out ? lab_out : lab_own.get();
    const int64_t nc = components_device(n, rp, ix, lab, root.get(), s);
    rcm_order_labeled(n, rp, ix, lab, root.get(), nc, perm, iperm, s);
    return nc;
}

void rcm_order_labeled(int64_t n, const int64_t* rp, const int32_t* ix, const int32_t* lab,
                       int32_t* start, int64_t nc, int32_t* perm, int32_t* iperm, hipStream_t s) {
    SI_CHECK(n >= 1 && n < (int64_t)INT32_MAX && nc >= 1 && nc <= n, kErrUnsupported,
             "rcm: n out of range");
    DevBuf<int32_t> queue(n), mark(n), lvl(n), cnt(n + 1), off(n + 1), maxlev(nc);
    DevBuf<unsigned long long> key(n), skey(n), skey2(n), best(nc);
    DevBuf<int> big(1);
    size_t tb_scan = 0;
    SI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb_scan, cnt.get(), off.get(), (int)(n + 1), s));
    DevBuf<char> tscan(tb_scan);
    DevBuf<char> tsort;
    hipLaunchKernelGGL(k_fill_u64, grid_of(n), dim3(256), 0, s, key.get(), n,
                       (unsigned long long)kNoClaim);
    hipLaunchKernelGGL(k_fill_i32, grid_of(n), dim3(256), 0, s, mark.get(), n, 0);
    SI_LAUNCH_CHECK();
    struct Readback {
        int32_t total;
        int big;
    };
    Readback rb{};
    for (int sweep = 0; sweep < 3; ++sweep) {
        const int32_t epoch = sweep + 1;
        const int sorted = sweep == 2 ? 1 : 0;
        hipLaunchKernelGGL(k_fill_i32, grid_of(nc), dim3(256), 0, s, maxlev.get(), nc, 0);
        hipLaunchKernelGGL(k_rcm_start, grid_of(nc), dim3(256), 0, s, (int32_t)nc, start,
                           queue.get(), mark.get(), lvl.get(), epoch);
        SI_LAUNCH_CHECK();
        int64_t ls = 0, le = nc;
        int32_t level = 0;
        while (le > ls) {
            const int64_t m = le - ls;
            ++level;
            hipLaunchKernelGGL(k_rcm_claim, grid_of(m * 64), dim3(256), 0, s, ls, m, rp, ix,
                               queue.get(), mark.get(), epoch, key.get());
            hipLaunchKernelGGL(k_rcm_count, grid_of((m + 1) * 64), dim3(256), 0, s, ls, m, rp, ix,
                               queue.get(), mark.get(), epoch, key.get(), cnt.get());
            SI_LAUNCH_CHECK();
            size_t tbl = tb_scan;
            SI_HIP(hipcub::DeviceScan::ExclusiveSum(tscan.get(), tbl, cnt.get(), off.get(),
                                                    (int)(m + 1), s));
            SI_HIP(hipMemsetAsync(big.get(), 0, sizeof(int), s));
            hipLaunchKernelGGL(k_rcm_emit, grid_of(m * 64), dim3(256), 0, s, ls, le, m, rp, ix,
                               queue.get(), mark.get(), epoch, key.get(), off.get(), lvl.get(),
                               level, lab, maxlev.get(), sorted, skey.get(), big.get());
            SI_LAUNCH_CHECK();
            SI_HIP(hipMemcpyAsync(&rb.total, off.get() + m, sizeof(int32_t), hipMemcpyDeviceToHost, s));
            SI_HIP(hipMemcpyAsync(&rb.big, big.get(), sizeof(int), hipMemcpyDeviceToHost, s));
            SI_HIP(hipStreamSynchronize(s));
            SI_CHECK(rb.total >= 0 && le + rb.total <= n, kErrInternal, "rcm: queue overflow");
            if (rb.big) {
                // a group above 64 nodes: sort every group of the level by (degree, id)
                size_t tb = 0;
                SI_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(
                    nullptr, tb, skey.get(), skey2.get(), rb.total, (int)m, off.get(), off.get() + 1,
                    0, 64, s));
                if (tb > tsort.n) tsort.alloc(tb);
                SI_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(
                    tsort.get(), tb, skey.get(), skey2.get(), rb.total, (int)m, off.get(),
                    off.get() + 1, 0, 64, s));
                hipLaunchKernelGGL(k_rcm_scatter, grid_of(rb.total), dim3(256), 0, s, le,
                                   (int64_t)rb.total, skey2.get(), queue.get());
                SI_LAUNCH_CHECK();
            }
            ls = le;
            le += rb.total;
        }
        SI_CHECK(le == n, kErrInternal, "rcm: a sweep did not reach every node");
        if (sweep < 2) {
            hipLaunchKernelGGL(k_fill_u64, grid_of(nc), dim3(256), 0, s, best.get(), nc,
                               (unsigned long long)kNoClaim);
            hipLaunchKernelGGL(k_rcm_last_level, grid_of(n), dim3(256), 0, s, n, rp, queue.get(),
                               lvl.get(), lab, maxlev.get(), best.get());
            hipLaunchKernelGGL(k_rcm_next_start, grid_of(nc), dim3(256), 0, s, (int32_t)nc,
                               best.get(), queue.get(), start);
            SI_LAUNCH_CHECK();
        }
    }
    // the global queue holds level L of every component before level L + 1:
    // a stable sort by component label gives the per-component orders in
    // component order
    const int32_t* order = queue.get();
    DevBuf<int32_t> ql, ql2, sorted_q;
    if (nc > 1) {
        ql.alloc(n);
        ql2.alloc(n);
        sorted_q.alloc(n);
        hipLaunchKernelGGL(k_queue_labels, grid_of(n), dim3(256), 0, s, n, queue.get(), lab,
                           ql.get());
        int bits = 1;
        while (bits < 31 && ((int64_t)1 << bits) < nc) ++bits;
        size_t tb = 0;
        SI_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, ql.get(), ql2.get(), queue.get(),
                                                  sorted_q.get(), (int)n, 0, bits, s));
        DevBuf<char> t(tb);
        SI_HIP(hipcub::DeviceRadixSort::SortPairs(t.get(), tb, ql.get(), ql2.get(), queue.get(),
                                                  sorted_q.get(), (int)n, 0, bits, s));
        SI_HIP(hipStreamSynchronize(s));
        order = sorted_q.get();
    }
    hipLaunchKernelGGL(k_rcm_reverse, grid_of(n), dim3(256), 0, s, n, order, perm, iperm);
    SI_LAUNCH_CHECK();
    SI_HIP(hipStreamSynchronize(s));
}

}  // namespace si
