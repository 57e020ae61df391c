"""CPU checks of the locality order that si_laplacian_build applies (reverse
Cuthill-McKee on host threads, spectral-init_amd/csrc/reorder.cpp), through the
host-only seam si_test_rcm_order: the threaded level expansion returns the
sequential queue's order exactly, on graphs whose BFS levels are large enough
to take the threaded path, including unsorted rows with repeated columns and
several components; a Python restatement of the sequential algorithm pins the
small cases; a ring's order has bandwidth 2."""
import numpy as np
import pytest
import scipy.sparse as sp

import spectral_mi355x as si


def random_graph(n, k, seed, dup=False):
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n), k)
    cols = rng.integers(0, n, n * k)
    a = sp.coo_matrix((np.ones(n * k), (rows, cols)), shape=(n, n)).tocsr()
    a = (a + a.T).tocsr()
    a.sum_duplicates()
    ip, ix = a.indptr.astype(np.int64), a.indices.astype(np.int32)
    if dup:  # shuffle every row and repeat its first column
        out_ix, out_ip = [], [0]
        for i in range(n):
            r = ix[ip[i]:ip[i + 1]].copy()
            rng.shuffle(r)
            if r.size:
                r = np.concatenate([r, r[:1]])
            out_ix.append(r)
            out_ip.append(out_ip[-1] + r.size)
        ip, ix = np.array(out_ip, np.int64), np.concatenate(out_ix).astype(np.int32)
    return ip, ix


def rcm_reference(ip, ix):
    """The sequential Cuthill-McKee queue (two unsorted sweeps from each
    component's first node for a pseudo-peripheral start), reversed."""
    n = ip.size - 1
    deg = np.diff(ip)

    def bfs(start, sort):
        seen = {start}
        order = [start]
        ls, le = 0, 1
        while True:
            for q in range(ls, le):
                v = order[q]
                nb = []
                for u in ix[ip[v]:ip[v + 1]]:
                    u = int(u)
                    if u not in seen:
                        seen.add(u)
                        nb.append(u)
                if sort:
                    nb.sort(key=lambda u: (deg[u], u))
                order.extend(nb)
            if len(order) == le:
                break
            ls, le = le, len(order)
        last = order[ls:le]
        best = last[0]
        for v in last:
            if deg[v] < deg[best]:
                best = v
        return order, best

    done = np.zeros(n, bool)
    order = []
    for s0 in range(n):
        if done[s0]:
            continue
        start = s0
        for _ in range(2):
            _, start = bfs(start, False)
        comp, _ = bfs(start, True)
        done[comp] = True
        order.extend(comp)
    return np.array(order[::-1], np.int32)


@pytest.mark.parametrize("dup", [False, True])
def test_rcm_matches_sequential_reference(dup):
    ip, ix = random_graph(3000, 3, 7, dup)
    # a second component
    ip2, ix2 = random_graph(500, 2, 8, dup)
    ip = np.concatenate([ip, ip[-1] + ip2[1:]])
    ix = np.concatenate([ix, ix2 + 3000]).astype(np.int32)
    ref = rcm_reference(ip, ix)
    for t in (1, 4):
        np.testing.assert_array_equal(si.rcm_order(ip, ix, t), ref)


@pytest.mark.parametrize("dup", [False, True])
def test_threaded_levels_give_the_sequential_order(dup):
    # levels of tens of thousands of nodes take the threaded expansion
    ip, ix = random_graph(200_000, 4, 11, dup)
    p1 = si.rcm_order(ip, ix, 1)
    assert np.array_equal(np.sort(p1), np.arange(ip.size - 1))
    for t in (2, 7, 16):
        np.testing.assert_array_equal(si.rcm_order(ip, ix, t), p1)


def test_ring_order_has_bandwidth_two():
    n = 40_000
    i = np.arange(n)
    a = sp.coo_matrix((np.ones(2 * n), (np.r_[i, i], np.r_[(i + 1) % n, (i - 1) % n])),
                      shape=(n, n)).tocsr()
    perm = si.rcm_order(a.indptr, a.indices, 8)
    iperm = np.empty(n, np.int64)
    iperm[perm] = np.arange(n)
    b = a.tocoo()
    assert np.abs(iperm[b.row] - iperm[b.col]).max() <= 2
