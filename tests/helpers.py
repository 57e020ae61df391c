"""Closed-form graphs and subspace metrics shared by the tests."""
from __future__ import annotations

import numpy as np


def ring_graph(n: int):
    """C_n adjacency, unit weights, sorted columns (tests/common/mod.rs:191-207)."""
    indptr = np.arange(0, 2 * n + 1, 2, dtype=np.int64)
    idx = np.empty(2 * n, np.int32)
    for i in range(n):
        a, b = (i - 1) % n, (i + 1) % n
        idx[2 * i], idx[2 * i + 1] = min(a, b), max(a, b)
    return indptr, idx, np.ones(2 * n, np.float32)


def path_graph(n: int):
    """P_n adjacency, unit weights (tests/common/mod.rs:211-228)."""
    rows = []
    for i in range(n):
        nb = [j for j in (i - 1, i + 1) if 0 <= j < n]
        rows.append(nb)
    indptr = np.zeros(n + 1, np.int64)
    indptr[1:] = np.cumsum([len(r) for r in rows])
    idx = np.array([j for r in rows for j in r], np.int32)
    return indptr, idx, np.ones(idx.size, np.float32)


def ring_eigenvalues(n: int, k: int):
    """Normalized Laplacian of C_n: 1 - cos(2 pi j / n), ascending (with multiplicity)."""
    lam = np.sort(1.0 - np.cos(2.0 * np.pi * np.arange(n) / n))
    return lam[:k]


def random_symmetric_graph(n: int, avg_deg: int, seed: int):
    """Random symmetric weighted graph with sorted CSR rows and no self loops."""
    rng = np.random.default_rng(seed)
    m = n * avg_deg // 2
    r = rng.integers(0, n, m)
    c = rng.integers(0, n, m)
    keep = r != c
    r, c = r[keep], c[keep]
    w = rng.uniform(0.05, 1.0, r.size).astype(np.float32)
    import scipy.sparse as sp
    a = sp.coo_matrix((w, (r, c)), shape=(n, n)).tocsr()
    a = a.maximum(a.T).tocsr()
    # make sure the graph is connected: add a ring with small weight
    ring = sp.coo_matrix((np.full(n, 0.01, np.float32), (np.arange(n), (np.arange(n) + 1) % n)),
                         shape=(n, n))
    a = (a + ring + ring.T).tocsr()
    a.sort_indices()
    a.data = a.data.astype(np.float32)
    return a.indptr.astype(np.int64), a.indices.astype(np.int32), a.data


def csr_to_dense(indptr, indices, data, n):
    d = np.zeros((n, n))
    for i in range(n):
        for p in range(indptr[i], indptr[i + 1]):
            d[i, indices[p]] += data[p]
    return d


def orthonormal(x):
    q, _ = np.linalg.qr(x)
    return q


def principal_angles(a, b):
    """Principal angles (radians, descending) between span(a) and span(b).

    Angles below pi/4 come from the sines (singular values of the part of
    span(b) outside span(a)), the rest from the cosines: arccos of a cosine
    one ulp below 1 is already 1.49e-8, so cosines alone cannot resolve small
    angles (Knyazev & Argentati, SIAM J. Sci. Comput. 23(6), 2002)."""
    qa, qb = orthonormal(np.asarray(a, float)), orthonormal(np.asarray(b, float))
    c = qa.T @ qb
    cos = np.clip(np.linalg.svd(c, compute_uv=False), -1.0, 1.0)
    sin = np.clip(np.linalg.svd(qb - qa @ c, compute_uv=False), 0.0, 1.0)
    big = np.arccos(cos)
    big = big[big >= np.pi / 4]
    small = np.arcsin(sin)
    small = small[small < np.pi / 4]
    k = min(qa.shape[1], qb.shape[1])
    ang = np.sort(np.concatenate([big, small]))[::-1]
    return ang[:k]


def max_subspace_angle(a, b):
    return float(principal_angles(a, b)[0])


def residuals_dense(lmat, evals, evecs):
    r = lmat @ evecs - evecs * evals[None, :]
    return np.linalg.norm(r, axis=0) / np.maximum(np.linalg.norm(evecs, axis=0), 1e-300)


def chained_random_graphs(nb: int, blocks: int, seed: int, links: int = 20, link_w: float = 0.5):
    """`blocks` random graphs of nb nodes (random_symmetric_graph, average degree
    8) joined in a chain by `links` edges of weight link_w between neighbouring
    blocks: no narrow band under any ordering (each block is an expander), and
    blocks - 1 small eigenvalues well below the rest of the spectrum."""
    import scipy.sparse as sp
    mats = []
    for b in range(blocks):
        ip, ix, w = random_symmetric_graph(nb, 8, seed + b)
        mats.append(sp.csr_matrix((w, ix, ip), shape=(nb, nb)))
    a = sp.block_diag(mats).tocoo()
    rng = np.random.default_rng(seed)
    r, c, v = [a.row], [a.col], [a.data]
    for b in range(blocks - 1):
        i = b * nb + rng.integers(0, nb, links)
        j = (b + 1) * nb + rng.integers(0, nb, links)
        r += [i, j]
        c += [j, i]
        v += [np.full(links, link_w, np.float32)] * 2
    n = nb * blocks
    m = sp.coo_matrix((np.concatenate(v), (np.concatenate(r), np.concatenate(c))), shape=(n, n))
    m = m.tocsr()
    m.sum_duplicates()
    m.sort_indices()
    return m.indptr.astype(np.int64), m.indices.astype(np.int32), m.data.astype(np.float32)
