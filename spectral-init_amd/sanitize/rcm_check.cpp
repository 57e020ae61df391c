// rcm_check.cpp -- host-only driver for the sanitizer builds of the locality
// order (reorder.cpp): AddressSanitizer + UndefinedBehaviorSanitizer, and
// ThreadSanitizer for the threaded level expansion (atomic-minimum parents).
// It orders a set of graphs on 1 and on 16 threads, checks that every result
// is a permutation and that the thread count does not change it, and exits
// non-zero on the first mismatch; the sanitizers abort on any report.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../csrc/si_device.hpp"

namespace {

struct Csr {
    int64_t n = 0;
    std::vector<int64_t> ip;
    std::vector<int32_t> ix;
};

// symmetric random graph, k random neighbours per node before symmetrising;
// shuffled rows, and with dup the first column of every row repeated
Csr random_graph(int64_t n, int k, uint64_t seed, bool dup, int64_t components) {
    std::mt19937_64 rng(seed);
    std::vector<std::vector<int32_t>> adj(n);
    const int64_t per = n / components;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t c = std::min<int64_t>(i / per, components - 1);
        const int64_t lo = c * per, hi = c == components - 1 ? n : lo + per;
        for (int t = 0; t < k; ++t) {
            const int64_t j = lo + (int64_t)(rng() % (uint64_t)(hi - lo));
            if (j == i) continue;
            adj[i].push_back((int32_t)j);
            adj[j].push_back((int32_t)i);
        }
    }
    Csr g;
    g.n = n;
    g.ip.assign(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        std::shuffle(adj[i].begin(), adj[i].end(), rng);
        if (dup && !adj[i].empty()) adj[i].push_back(adj[i].front());
        g.ip[i + 1] = g.ip[i] + (int64_t)adj[i].size();
        g.ix.insert(g.ix.end(), adj[i].begin(), adj[i].end());
    }
    return g;
}

Csr ring(int64_t n) {
    Csr g;
    g.n = n;
    g.ip.assign(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t a = (i + n - 1) % n, b = (i + 1) % n;
        g.ix.push_back((int32_t)std::min(a, b));
        g.ix.push_back((int32_t)std::max(a, b));
        g.ip[i + 1] = g.ip[i] + 2;
    }
    return g;
}

bool is_permutation(const std::vector<int32_t>& p, int64_t n) {
    if ((int64_t)p.size() != n) return false;
    std::vector<char> seen(n, 0);
    for (int32_t v : p) {
        if (v < 0 || v >= n || seen[v]) return false;
        seen[v] = 1;
    }
    return true;
}

}  // namespace

int main() {
    struct Case {
        const char* name;
        Csr g;
    };
    std::vector<Case> cases;
    cases.push_back({"random60k", random_graph(60000, 12, 1, false, 1)});
    cases.push_back({"random60k_dup", random_graph(60000, 12, 2, true, 1)});
    cases.push_back({"random80k_4comp", random_graph(80000, 10, 3, false, 4)});
    cases.push_back({"ring5000", ring(5000)});
    cases.push_back({"tiny", random_graph(7, 2, 4, true, 2)});
    for (const Case& c : cases) {
        const std::vector<int32_t> p1 = si::rcm_order(c.g.n, c.g.ip.data(), c.g.ix.data(), 1);
        const std::vector<int32_t> p16 = si::rcm_order(c.g.n, c.g.ip.data(), c.g.ix.data(), 16);
        if (!is_permutation(p1, c.g.n) || !is_permutation(p16, c.g.n)) {
            std::printf("%s: not a permutation\n", c.name);
            return 1;
        }
        if (p1 != p16) {
            std::printf("%s: 1 and 16 threads give different orders\n", c.name);
            return 1;
        }
        std::printf("%s: n=%lld ok\n", c.name, (long long)c.g.n);
    }
    std::printf("rcm_check: all cases ok\n");
    return 0;
}
