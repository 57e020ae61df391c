// reorder.cpp -- node reordering for gather locality of the block SpMM.
//
// A k-NN graph of shuffled points scatters every row's neighbours across the
// whole x block, so each gathered x row misses in the XCD's L2.  Reverse
// Cuthill-McKee (BFS from a pseudo-peripheral node, neighbours visited by
// increasing degree, order reversed) places neighbours within a narrow band
// of row indices, so the rows an XCD streams gather from a compact, L2-resident
// window of x.  The permutation is applied once when the Laplacian is built;
// the solvers run in the permuted order and every public entry point maps
// vectors back.
//
// The BFS runs level by level on host threads and yields exactly the order of
// the sequential queue: a node of the next level belongs to the first node of
// the current level (in queue order) adjacent to it, found as an atomic
// minimum of parent positions; each parent's new nodes keep their adjacency
// order (or are sorted by degree, then id), and the groups follow in parent
// order.  Levels below kParLevel nodes run sequentially.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <thread>
#include <vector>

#include "si_host.hpp"

namespace si {

namespace {

constexpr size_t kParLevel = 16384;  // smallest level expanded on several threads

// slot = min(slot, q), atomically (compare-exchange loop on the int32)
inline void atomic_min_i32(int32_t& slot, int32_t q) {
    int32_t cur = __atomic_load_n(&slot, __ATOMIC_RELAXED);
    while (q < cur &&
           !__atomic_compare_exchange_n(&slot, &cur, q, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
}

struct Bfs {
    const int64_t* ip;
    const int32_t* ix;
    std::vector<int>& mark;          // mark[v] == epoch: visited in this BFS
    std::vector<int32_t>& parent;    // INT32_MAX outside the level being expanded
    int nth;

    int64_t deg(int64_t v) const { return ip[v + 1] - ip[v]; }
    bool by_degree(int64_t x, int64_t y) const {
        const int64_t dx = deg(x), dy = deg(y);
        return dx != dy ? dx < dy : x < y;
    }

    // Appends level L = order[ls, le) 's children to order (sequential form).
    void expand_seq(std::vector<int64_t>& order, size_t ls, size_t le, int epoch, bool sorted,
                    std::vector<int64_t>& nb) {
        for (size_t q = ls; q < le; ++q) {
            const int64_t v = order[q];
            nb.clear();
            for (int64_t p = ip[v]; p < ip[v + 1]; ++p) {
                const int64_t u = ix[p];
                if (mark[u] != epoch) {
                    mark[u] = epoch;
                    nb.push_back(u);
                }
            }
            if (sorted)
                std::sort(nb.begin(), nb.end(), [&](int64_t x, int64_t y) { return by_degree(x, y); });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }

    // The same level expansion on nth threads, identical result.
    void expand_par(std::vector<int64_t>& order, size_t ls, size_t le, int epoch, bool sorted) {
        const size_t m = le - ls;
        const int T = nth;
        // 1. parent[u] = min position in the level of a neighbour (unvisited u)
        auto claim = [&](size_t a, size_t b) {
            for (size_t q = a; q < b; ++q) {
                const int64_t v = order[ls + q];
                for (int64_t p = ip[v]; p < ip[v + 1]; ++p) {
                    const int64_t u = ix[p];
                    if (mark[u] == epoch) continue;
                    atomic_min_i32(parent[u], (int32_t)q);
                }
            }
        };
        run(claim, m, T);
        // 2. each chunk gathers its parents' groups in position order
        std::vector<std::vector<int64_t>> part(T);
        auto gather = [&](int t, size_t a, size_t b) {
            std::vector<int64_t>& out = part[t];
            for (size_t q = a; q < b; ++q) {
                const int64_t v = order[ls + q];
                const size_t g0 = out.size();
                for (int64_t p = ip[v]; p < ip[v + 1]; ++p) {
                    const int64_t u = ix[p];
                    // the owner emits u once (a row may repeat a column) and
                    // retires the claim; other chunks only compare it with
                    // their own positions
                    if (mark[u] != epoch &&
                        __atomic_load_n(&parent[u], __ATOMIC_RELAXED) == (int32_t)q) {
                        out.push_back(u);
                        __atomic_store_n(&parent[u], (int32_t)-1, __ATOMIC_RELAXED);
                    }
                }
                if (sorted)
                    std::sort(out.begin() + g0, out.end(),
                              [&](int64_t x, int64_t y) { return by_degree(x, y); });
            }
        };
        run_indexed(gather, m, T);
        // 3. concatenate, mark visited, reset parents
        for (int t = 0; t < T; ++t) {
            for (int64_t u : part[t]) {
                mark[u] = epoch;
                parent[u] = INT32_MAX;
            }
            order.insert(order.end(), part[t].begin(), part[t].end());
        }
    }

    template <typename F>
    static void run(F&& f, size_t m, int T) {
        std::vector<std::thread> pool;
        for (int t = 1; t < T; ++t) pool.emplace_back(f, m * t / T, m * (t + 1) / T);
        f(0, m / T);
        for (auto& th : pool) th.join();
    }
    template <typename F>
    static void run_indexed(F&& f, size_t m, int T) {
        std::vector<std::thread> pool;
        for (int t = 1; t < T; ++t) pool.emplace_back(f, t, m * t / T, m * (t + 1) / T);
        f(0, 0, m / T);
        for (auto& th : pool) th.join();
    }

    // BFS over one component from `start`, appending to `order`; returns the
    // first minimum-degree node (queue order) of the last level.
    int64_t levels(int64_t start, int epoch, std::vector<int64_t>& order, bool sorted) {
        const size_t base = order.size();
        order.push_back(start);
        mark[start] = epoch;
        size_t ls = base, le = base + 1;
        std::vector<int64_t> nb;
        for (;;) {
            if (nth > 1 && le - ls >= kParLevel) expand_par(order, ls, le, epoch, sorted);
            else expand_seq(order, ls, le, epoch, sorted, nb);
            if (order.size() == le) break;
            ls = le;
            le = order.size();
        }
        int64_t best = order[ls];
        for (size_t q = ls; q < le; ++q)
            if (deg(order[q]) < deg(best)) best = order[q];
        return best;
    }
};

}  // namespace

// Reverse Cuthill-McKee order (perm[new] = old) of a symmetric CSR pattern on
// up to `threads` host threads (0: up to 16, as the hardware allows); the
// result does not depend on the thread count.
std::vector<int32_t> rcm_order(int64_t n, const int64_t* ip, const int32_t* ix, int threads) {
    std::vector<int> mark(n, 0);
    std::vector<int32_t> parent(n, INT32_MAX);
    if (threads <= 0)
        threads = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    Bfs bfs{ip, ix, mark, parent, threads};
    std::vector<int64_t> order, scratch;
    order.reserve(n);
    int epoch = 1;
    std::vector<char> done(n, 0);
    for (int64_t s0 = 0; s0 < n; ++s0) {
        if (done[s0]) continue;
        // pseudo-peripheral start: two sweeps from s0
        int64_t start = s0;
        for (int sweep = 0; sweep < 2; ++sweep) {
            scratch.clear();
            ++epoch;
            start = bfs.levels(start, epoch, scratch, false);
        }
        ++epoch;
        const size_t b = order.size();
        bfs.levels(start, epoch, order, true);
        for (size_t q = b; q < order.size(); ++q) done[order[q]] = 1;
    }
    std::vector<int32_t> perm(n);
    for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)order[n - 1 - i];
    return perm;
}

}  // namespace si
