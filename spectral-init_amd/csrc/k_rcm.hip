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
// repeated columns.  Per level: claim; count (one wavefront per level node)
// with the exclusive scan of the counts inside contiguous chunks; the scan of
// the chunk sums, which also writes the end of the next level into the level
// bounds kept on the device; emit (ballot ranks in adjacency order; in the
// last sweep groups of up to kGroupLds new nodes are ranked by (degree, id) in
// LDS, larger groups by a segmented radix sort).  The host enqueues levels in
// batches and reads the bounds back once per batch -- a level past the end of
// a sweep is empty and its kernels return at once -- except in the last sweep
// of a graph with a row longer than kGroupLds, which reads back every level.
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

// hooking: f[f[u]] and f[u] to min(f[f[v]]) over the neighbours v of u.
// kHookLanes lanes per node share its row; a workgroup takes 256 / kHookLanes
// consecutive nodes at a time, grid-strided.  The root entry f[f[u]] is shared
// by many nodes: it is read before its atomic, which only a lower value needs.
constexpr int kHookLanes = 16;

__global__ __launch_bounds__(256) void k_cc_hook(int64_t n, const int64_t* __restrict__ rp,
                                                 const int32_t* __restrict__ ix,
                                                 int32_t* __restrict__ f, int* __restrict__ changed) {
    constexpr int64_t kPerBlock = 256 / kHookLanes;
    const int sub = threadIdx.x & (kHookLanes - 1);
    bool ch = false;
    for (int64_t base = (int64_t)blockIdx.x * kPerBlock; base < n;
         base += (int64_t)gridDim.x * kPerBlock) {
        const int64_t u = base + threadIdx.x / kHookLanes;
        int32_t best = INT32_MAX;
        if (u < n) {
            const int64_t e = rp[u + 1];
            for (int64_t p = rp[u] + sub; p < e; p += kHookLanes) best = min(best, f[f[ix[p]]]);
        }
        for (int o = kHookLanes / 2; o > 0; o >>= 1) best = min(best, __shfl_xor(best, o));
        if (u < n && sub == 0 && best != INT32_MAX) {
            const int32_t fu = f[u];
            if (best < fu) {
                if (f[fu] > best) ch |= atomicMin(&f[fu], best) > best;
                ch |= atomicMin(&f[u], best) > best;
            }
        }
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

// Workgroups of the level kernels.  count splits a level into kLevelGrid
// contiguous chunks; k_rcm_parts scans the chunk sums with one workgroup of
// kLevelGrid / 2 threads, two sums each.
constexpr int kLevelGrid = 2048;
// In the sorted sweep a group of up to this many new nodes is ranked in LDS.
constexpr int kGroupLds = 1024;
// Levels enqueued between two readbacks of the level bounds.
constexpr int kLevelBatch = 16;

// The state the level kernels share.  Level L occupies queue[bnd[L] ..
// bnd[L + 1]); expanding it writes bnd[L + 2].
struct Level {
    const int64_t* rp;
    const int32_t* ix;
    int32_t* queue;
    int32_t* mark;
    unsigned long long* key;
    int32_t* bnd;
    int32_t* cnt;   // new nodes owned by each node of the level
    int32_t* loc;   // their exclusive sum within the node's chunk
    int32_t* part;  // chunk sums, then their exclusive sums
    int32_t* lvl;
    const int32_t* lab;
    int32_t* maxlev;
    unsigned long long* skey;
    int* big;
    int32_t epoch;
};

__device__ __forceinline__ int64_t level_chunk(int64_t m) {
    return (m + kLevelGrid - 1) / kLevelGrid;
}

__device__ __forceinline__ int32_t wave_incl_scan(int32_t x, int lane) {
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    return x;
}

// level 0 of a sweep: queue[c] = start[c]
__global__ void k_rcm_start(int32_t nc, const int32_t* __restrict__ start, int32_t* __restrict__ queue,
                            int32_t* __restrict__ mark, int32_t* __restrict__ lvl,
                            int32_t* __restrict__ bnd, int32_t epoch) {
    const int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c == 0) {
        bnd[0] = 0;
        bnd[1] = nc;
    }
    if (c >= nc) return;
    const int32_t v = start[c];
    queue[c] = v;
    mark[v] = epoch;
    lvl[c] = 0;
}

// claim: key[u] = min (position, entry) over the level's entries (q, p) with
// unvisited column u; one wavefront per level node
__global__ __launch_bounds__(256) void k_rcm_claim(Level A, int L) {
    const int64_t ls = A.bnd[L], m = A.bnd[L + 1] - ls;
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < m; j += waves) {
        const int64_t q = ls + j;
        const int32_t v = A.queue[q];
        const int64_t e = A.rp[v + 1];
        for (int64_t p = A.rp[v] + lane; p < e; p += 64) {
            const int32_t u = A.ix[p];
            if (A.mark[u] != A.epoch) atomicMin(&A.key[u], (unsigned long long)claim_key(q, p));
        }
    }
}

// count: cnt[j] = entries of level node j that own their column, and loc[j]
// their exclusive sum within the workgroup's chunk; part[b] = chunk b's total
__global__ __launch_bounds__(256) void k_rcm_count(Level A, int L) {
    __shared__ int32_t tc[256];
    __shared__ int32_t wsum[4];
    const int64_t ls = A.bnd[L], m = A.bnd[L + 1] - ls;
    const int64_t C = level_chunk(m);
    const int64_t b0 = min(m, (int64_t)blockIdx.x * C), b1 = min(m, b0 + C);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int32_t running = 0;
    for (int64_t t0 = b0; t0 < b1; t0 += 256) {
        const int nt = (int)min((int64_t)256, b1 - t0);
        for (int jj = w; jj < nt; jj += 4) {
            const int64_t q = ls + t0 + jj;
            const int32_t v = A.queue[q];
            const int64_t e = A.rp[v + 1];
            int c = 0;
            for (int64_t p = A.rp[v] + lane; p < e; p += 64) {
                const int32_t u = A.ix[p];
                c += (A.mark[u] != A.epoch && A.key[u] == claim_key(q, p)) ? 1 : 0;
            }
            for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
            if (lane == 0) tc[jj] = c;
        }
        __syncthreads();
        const int32_t x = (int)threadIdx.x < nt ? tc[threadIdx.x] : 0;
        const int32_t incl = wave_incl_scan(x, lane);
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        int32_t base = running;
        for (int i = 0; i < w; ++i) base += wsum[i];
        if ((int)threadIdx.x < nt) {
            A.cnt[t0 + threadIdx.x] = x;
            A.loc[t0 + threadIdx.x] = base + incl - x;
        }
        running += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();  // tc and wsum are reused by the next tile
    }
    if (threadIdx.x == 0) A.part[blockIdx.x] = running;
}

// exclusive scan of the kLevelGrid chunk sums (in place); the level's total
// gives the end of the next level
__global__ __launch_bounds__(kLevelGrid / 2) void k_rcm_parts(Level A, int L) {
    __shared__ int32_t wsum[kLevelGrid / 128];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int32_t a = A.part[2 * t], b = A.part[2 * t + 1];
    const int32_t x = a + b;
    const int32_t incl = wave_incl_scan(x, lane);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int32_t base = 0;
    for (int i = 0; i < w; ++i) base += wsum[i];
    const int32_t ex = base + incl - x;
    A.part[2 * t] = ex;
    A.part[2 * t + 1] = ex + a;
    if (t == kLevelGrid / 2 - 1) A.bnd[L + 2] = A.bnd[L + 1] + base + incl;
}

// emit: the owned columns of level node j go to the next level at offset
// part[chunk] + loc[j], in adjacency order; the sorted sweep ranks a group of
// <= kGroupLds by (degree, id) in LDS, and writes every group's keys to skey
// (a larger group, flagged in *big, is sorted afterwards by the host's
// segmented sort)
template <bool SORTED>
__global__ __launch_bounds__(256) void k_rcm_emit(Level A, int L) {
    __shared__ unsigned long long grp[SORTED ? 4 : 1][SORTED ? kGroupLds : 1];
    const int64_t ls = A.bnd[L], le = A.bnd[L + 1], m = le - ls;
    const int64_t C = level_chunk(m);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int32_t level = L + 1;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    int32_t seen = -1;  // the component whose maxlev this wave last set
    for (int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < m; j += waves) {
        const int32_t g = A.cnt[j];
        const int64_t q = ls + j;
        const int32_t v = A.queue[q];
        if (g == 0) continue;
        // the deepest level that received nodes, per component: every writer
        // of this launch stores the same level, and later levels are larger
        const int32_t c = A.lab[v];
        if (c != seen) {
            if (lane == 0) A.maxlev[c] = level;
            seen = c;
        }
        const int64_t o = (int64_t)A.part[j / C] + A.loc[j];
        const int64_t e = A.rp[v + 1];
        int32_t run = 0;
        for (int64_t p0 = A.rp[v]; p0 < e; p0 += 64) {
            const int64_t p = p0 + lane;
            int32_t u = -1;
            bool own = false;
            if (p < e) {
                u = A.ix[p];
                own = A.mark[u] != A.epoch && A.key[u] == claim_key(q, p);
            }
            const uint64_t bal = __ballot(own);
            if (own) {
                const int32_t r = run + __popcll(bal & ((1ull << lane) - 1ull));
                if (!SORTED) {
                    A.queue[le + o + r] = u;
                } else {
                    const unsigned long long dk =
                        ((uint64_t)(A.rp[u + 1] - A.rp[u]) << 32) | (uint32_t)u;
                    if (g <= kGroupLds) grp[w][r] = dk;
                    else A.skey[o + r] = dk;
                }
                A.lvl[le + o + r] = level;
                A.mark[u] = A.epoch;
                A.key[u] = kNoClaim;
            }
            run += __popcll(bal);
        }
        if (SORTED) {
            if (g <= kGroupLds) {
                __syncwarp();
                for (int i = lane; i < g; i += 64) {
                    const unsigned long long mine = grp[w][i];
                    int32_t r = 0;
                    for (int t = 0; t < g; ++t) r += grp[w][t] < mine ? 1 : 0;
                    A.queue[le + o + r] = (int32_t)(uint32_t)mine;
                    A.skey[o + r] = mine;
                }
                __syncwarp();  // the wave's next group reuses grp[w]
            } else if (lane == 0) {
                *A.big = 1;
            }
        }
    }
}

// begin offsets of every group of the level (off[m] = the level's total), for
// the segmented sort
__global__ void k_rcm_offsets(Level A, int L, int32_t* __restrict__ off) {
    const int64_t ls = A.bnd[L], m = A.bnd[L + 1] - ls;
    const int64_t C = level_chunk(m);
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m) off[j] = A.part[j / C] + A.loc[j];
    else if (j == m) off[m] = A.bnd[L + 2] - A.bnd[L + 1];
}

// after the segmented sort of a level: queue[le + i] = id of sorted key i
__global__ void k_rcm_scatter(int64_t le, int64_t cnt, const unsigned long long* __restrict__ sk,
                              int32_t* __restrict__ queue) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) queue[le + i] = (int32_t)(uint32_t)sk[i];
}

// the longest stored row (an upper bound of every new-node group)
__global__ __launch_bounds__(256) void k_max_row(int64_t n, const int64_t* __restrict__ rp,
                                                 int* __restrict__ out) {
    __shared__ int wmax[4];
    int len = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        len = max(len, (int)min(rp[i + 1] - rp[i], (int64_t)INT_MAX));
    for (int o = 32; o > 0; o >>= 1) len = max(len, __shfl_xor(len, o));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = len;
    __syncthreads();
    if (threadIdx.x == 0) {
        len = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
        if (len > 0) atomicMax(out, len);
    }
}

// first minimum-degree node (queue order) of each component's last level
__global__ void k_rcm_last_level(int64_t n, const int64_t* __restrict__ rp,
                                 const int32_t* __restrict__ queue, const int32_t* __restrict__ lvl,
                                 const int32_t* __restrict__ lab, const int32_t* __restrict__ maxlev,
                                 unsigned long long* __restrict__ best) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int32_t c = -1;
    unsigned long long key = kNoClaim;
    if (q < n) {
        const int32_t v = queue[q];
        c = lab[v];
        if (lvl[q] == maxlev[c]) key = ((uint64_t)(rp[v + 1] - rp[v]) << 32) | (uint64_t)q;
    }
    // the wave's candidates, one component at a time (the queue is ordered
    // by level and component, so a wave usually holds one)
    bool pend = key != kNoClaim;
    for (;;) {
        const uint64_t bal = __ballot(pend);
        if (bal == 0) break;
        const int leader = __ffsll((unsigned long long)bal) - 1;
        const int32_t cl = __shfl(c, leader);
        unsigned long long mine = (pend && c == cl) ? key : kNoClaim;
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long other = __shfl_xor(mine, o);
            mine = other < mine ? other : mine;
        }
        if (lane == leader) atomicMin(&best[cl], mine);
        if (c == cl) pend = false;
    }
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
    StreamBuf<int32_t> f(n, s), rk(n + 1, s);
    StreamBuf<int> flag(1, s);
    hipLaunchKernelGGL(k_cc_init, grid_of(n), dim3(256), 0, s, n, f.get());
    SI_LAUNCH_CHECK();
    // every changing round lowers some pointer, so n rounds bound the loop; in
    // practice it ends after O(log diameter) rounds
    int64_t rounds = 0;
    for (;;) {
        SI_CHECK(rounds++ <= n + 1, kErrInternal, "components: no fixed point");
        SI_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), s));
        hipLaunchKernelGGL(k_cc_hook,
                           dim3((unsigned)std::min<int64_t>(ceil_div(n * kHookLanes, 256), 4096)),
                           dim3(256), 0, s, n, rp, ix, f.get(), flag.get());
        hipLaunchKernelGGL(k_cc_shortcut, grid_of(n), dim3(256), 0, s, n, f.get(), flag.get());
        SI_LAUNCH_CHECK();
        int ch = 0;
        SI_HIP(hipMemcpyAsync(&ch, flag.get(), sizeof(int), hipMemcpyDeviceToHost, s));
        SI_HIP(hipStreamSynchronize(s));
        if (!ch) break;
    }
    StreamBuf<int32_t> isroot(n + 1, s);
    hipLaunchKernelGGL(k_cc_roots, grid_of(n + 1), dim3(256), 0, s, n, f.get(), isroot.get());
    size_t tb = 0;
    SI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, isroot.get(), rk.get(), (int)(n + 1), s));
    StreamBuf<char> tmp(tb, s);
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
    StreamBuf<int32_t> lab_own(lab_out ? 0 : n, s), root(n, s);
    int32_t* lab = lab_out ? lab_out : lab_own.get();
    const int64_t nc = components_device(n, rp, ix, lab, root.get(), s);
    rcm_order_labeled(n, rp, ix, lab, root.get(), nc, perm, iperm, s);
    return nc;
}

void rcm_order_labeled(int64_t n, const int64_t* rp, const int32_t* ix, const int32_t* lab,
                       int32_t* start, int64_t nc, int32_t* perm, int32_t* iperm, hipStream_t s) {
    SI_CHECK(n >= 1 && n < (int64_t)INT32_MAX && nc >= 1 && nc <= n, kErrUnsupported,
             "rcm: n out of range");
    StreamBuf<int32_t> queue(n, s), mark(n, s), lvl(n, s), cnt(n, s), loc(n, s),
        part(kLevelGrid, s), bnd(n + 2, s), maxlev(nc, s);
    StreamBuf<unsigned long long> key(n, s), skey(n, s), best(nc, s);
    StreamBuf<int> flags(2, s);  // [0] a group above kGroupLds, [1] the longest row
    hipLaunchKernelGGL(k_fill_u64, grid_of(n), dim3(256), 0, s, key.get(), n,
                       (unsigned long long)kNoClaim);
    hipLaunchKernelGGL(k_fill_i32, grid_of(n), dim3(256), 0, s, mark.get(), n, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, s, flags.get(), 2, 0);
    hipLaunchKernelGGL(k_max_row, dim3((unsigned)std::min<int64_t>(ceil_div(n, 256), 1024)), dim3(256),
                       0, s, n, rp, flags.get() + 1);
    SI_LAUNCH_CHECK();
    int max_row = 0;
    SI_HIP(hipMemcpyAsync(&max_row, flags.get() + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    SI_HIP(hipStreamSynchronize(s));
    // a group of new nodes is at most a row long: only a graph with a row
    // above kGroupLds can need the segmented sort, and its sorted sweep then
    // reads back every level
    const bool long_rows = max_row > kGroupLds;
    StreamBuf<int32_t> off(long_rows ? n + 1 : 0, s);
    StreamBuf<unsigned long long> skey2(long_rows ? n : 0, s);
    StreamBuf<char> tsort;
    Level A{rp,        ix,        queue.get(), mark.get(), key.get(),    bnd.get(),   cnt.get(),
            loc.get(), part.get(), lvl.get(),  lab,        maxlev.get(), skey.get(), flags.get(),
            0};
    int64_t depth = 0;  // levels of the previous sweep
    for (int sweep = 0; sweep < 3; ++sweep) {
        A.epoch = sweep + 1;
        const bool sorted = sweep == 2;
        const bool each = sorted && long_rows;
        hipLaunchKernelGGL(k_fill_i32, grid_of(nc), dim3(256), 0, s, maxlev.get(), nc, 0);
        hipLaunchKernelGGL(k_rcm_start, grid_of(nc), dim3(256), 0, s, (int32_t)nc, start,
                           queue.get(), mark.get(), lvl.get(), bnd.get(), A.epoch);
        SI_LAUNCH_CHECK();
        // Levels are enqueued in batches (the first as deep as the previous
        // sweep); a level past the end of the sweep is empty and its kernels
        // return at once.  Level L is expanded by its launches with argument
        // L; at most n levels are non-empty.
        int64_t L = 0;
        int64_t batch = depth > 0 ? depth : kLevelBatch;
        int32_t b3[3] = {0, 0, 0};
        for (;;) {
            const int64_t nb = std::min<int64_t>(each ? 1 : batch, n - L);
            SI_CHECK(nb > 0, kErrInternal, "rcm: more levels than nodes");
            if (each) SI_HIP(hipMemsetAsync(flags.get(), 0, sizeof(int), s));
            for (int64_t i = 0; i < nb; ++i, ++L) {
                hipLaunchKernelGGL(k_rcm_claim, dim3(kLevelGrid), dim3(256), 0, s, A, (int)L);
                hipLaunchKernelGGL(k_rcm_count, dim3(kLevelGrid), dim3(256), 0, s, A, (int)L);
                hipLaunchKernelGGL(k_rcm_parts, dim3(1), dim3(kLevelGrid / 2), 0, s, A, (int)L);
                if (sorted)
                    hipLaunchKernelGGL(k_rcm_emit<true>, dim3(kLevelGrid), dim3(256), 0, s, A, (int)L);
                else
                    hipLaunchKernelGGL(k_rcm_emit<false>, dim3(kLevelGrid), dim3(256), 0, s, A, (int)L);
                SI_LAUNCH_CHECK();
            }
            // bounds of the last expanded level (L - 1) and of the new level L
            int big = 0;
            SI_HIP(hipMemcpyAsync(b3, bnd.get() + L - 1, sizeof(b3), hipMemcpyDeviceToHost, s));
            if (each) SI_HIP(hipMemcpyAsync(&big, flags.get(), sizeof(int), hipMemcpyDeviceToHost, s));
            SI_HIP(hipStreamSynchronize(s));
            SI_CHECK(0 <= b3[0] && b3[0] <= b3[1] && b3[1] <= b3[2] && b3[2] <= n, kErrInternal,
                     "rcm: queue overflow");
            if (big) {
                // a group above kGroupLds nodes: sort every group of level L by
                // (degree, id)
                const int64_t m = b3[1] - b3[0], total = b3[2] - b3[1];
                hipLaunchKernelGGL(k_rcm_offsets, grid_of(m + 1), dim3(256), 0, s, A, (int)(L - 1),
                                   off.get());
                SI_LAUNCH_CHECK();
                size_t tb = 0;
                SI_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(
                    nullptr, tb, skey.get(), skey2.get(), (int)total, (int)m, off.get(),
                    off.get() + 1, 0, 64, s));
                if (tb > tsort.n) tsort.alloc(tb, s);
                SI_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(
                    tsort.get(), tb, skey.get(), skey2.get(), (int)total, (int)m, off.get(),
                    off.get() + 1, 0, 64, s));
                hipLaunchKernelGGL(k_rcm_scatter, grid_of(total), dim3(256), 0, s, (int64_t)b3[1],
                                   total, skey2.get(), queue.get());
                SI_LAUNCH_CHECK();
            }
            if (b3[2] == b3[1]) break;  // level L is empty
            batch = kLevelBatch;
        }
        SI_CHECK(b3[1] == n, kErrInternal, "rcm: a sweep did not reach every node");
        if (sweep < 2) {
            // the sweep's depth: its first empty level
            std::vector<int32_t> hb(L + 1);
            to_host(hb.data(), bnd.get(), (size_t)(L + 1), s);
            depth = 1;
            while (depth < L && hb[depth] < hb[depth + 1]) ++depth;
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
    StreamBuf<int32_t> ql, ql2, sorted_q;
    if (nc > 1) {
        ql.alloc(n, s);
        ql2.alloc(n, s);
        sorted_q.alloc(n, s);
        hipLaunchKernelGGL(k_queue_labels, grid_of(n), dim3(256), 0, s, n, queue.get(), lab,
                           ql.get());
        int bits = 1;
        while (bits < 31 && ((int64_t)1 << bits) < nc) ++bits;
        size_t tb = 0;
        SI_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, ql.get(), ql2.get(), queue.get(),
                                                  sorted_q.get(), (int)n, 0, bits, s));
        StreamBuf<char> t(tb, s);
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
