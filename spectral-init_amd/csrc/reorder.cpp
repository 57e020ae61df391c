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
#include <algorithm>
#include <cstdint>
#include <vector>

#include "si_host.hpp"

namespace si {

namespace {

// BFS over one component from `start`; fills `order` (appended) and returns a
// minimum-degree node of the last level.  mark[v] == epoch means visited.
int64_t bfs_levels(int64_t start, const int64_t* ip, const int32_t* ix, std::vector<int>& mark,
                   int epoch, std::vector<int64_t>& order, bool sort_by_degree) {
    const size_t base = order.size();
    order.push_back(start);
    mark[start] = epoch;
    size_t head = base;
    size_t level_start = base;
    size_t level_end = base + 1;
    std::vector<int64_t> nb;
    while (head < order.size()) {
        if (head == level_end) {
            level_start = level_end;
            level_end = order.size();
        }
        const int64_t v = order[head++];
        nb.clear();
        for (int64_t p = ip[v]; p < ip[v + 1]; ++p) {
            const int64_t u = ix[p];
            if (mark[u] != epoch) {
                mark[u] = epoch;
                nb.push_back(u);
            }
        }
        if (sort_by_degree)
            std::sort(nb.begin(), nb.end(), [&](int64_t x, int64_t y) {
                const int64_t dx = ip[x + 1] - ip[x], dy = ip[y + 1] - ip[y];
                return dx != dy ? dx < dy : x < y;
            });
        order.insert(order.end(), nb.begin(), nb.end());
    }
    // last level = [level_start', end): recompute as the nodes after the last
    // level boundary reached
    if (level_end < order.size()) level_start = level_end;
    int64_t best = order[level_start];
    for (size_t q = level_start; q < order.size(); ++q) {
        const int64_t v = order[q];
        if (ip[v + 1] - ip[v] < ip[best + 1] - ip[best]) best = v;
    }
    return best;
}

}  // namespace

// Reverse Cuthill-McKee order (perm[new] = old) of a symmetric CSR pattern.
std::vector<int32_t> rcm_order(int64_t n, const int64_t* ip, const int32_t* ix) {
    std::vector<int> mark(n, 0);
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
            start = bfs_levels(start, ip, ix, mark, epoch, scratch, false);
        }
        ++epoch;
        const size_t b = order.size();
        bfs_levels(start, ip, ix, mark, epoch, order, true);
        for (size_t q = b; q < order.size(); ++q) done[order[q]] = 1;
    }
    std::vector<int32_t> perm(n);
    for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)order[n - 1 - i];
    return perm;
}

}  // namespace si
