"""Synthetic fuzzy k-NN graphs: the inputs of the eigensolver path.

This is input preparation (the step before the reference's hot path), used by
the tests, the benchmark and the smoke check to produce seeded workloads:

  points   Gaussian mixtures (optionally with a uniform background fraction)
  kNN      exact k nearest neighbours, self excluded, Euclidean
  weights  UMAP fuzzy weights w_ij = exp(-(d_ij - rho_i)/sigma_i), rho_i the
           nearest-neighbour distance, sigma_i by binary search so that
           sum_j w_ij = log2(k)
  union    W = w + w^T - w o w^T (float32, bitwise symmetric), CSR with sorted
           columns, int64 row pointers, int32 columns

The float32 union formula is the fuzzy union of the reference's fixture
generator (tests/generate_fixtures.py:145-160).  The distance/top-k work runs
through torch (on the GPU when one is present): it is data preparation, not the
measured path.
"""
from __future__ import annotations

import math

import numpy as np


# ----------------------------------------------------------------- points
def gmm_points(n, dim, sizes, seed, center_scale=10.0, background_frac=0.0):
    """Gaussian mixture with unit-variance clusters; centers ~ N(0, center_scale^2 I).

    sizes: cluster proportions (normalised) or absolute sizes summing to the
    non-background count.  The background fraction is drawn uniformly from the
    bounding box of the centers.  Points are shuffled (no locality for free).
    """
    rng = np.random.default_rng(seed)
    n_bg = int(round(background_frac * n))
    n_fg = n - n_bg
    p = np.asarray(sizes, dtype=np.float64)
    p = p / p.sum()
    counts = np.floor(p * n_fg).astype(np.int64)
    counts[: n_fg - counts.sum()] += 1
    centers = rng.standard_normal((p.size, dim)) * center_scale
    x = np.empty((n, dim), np.float32)
    labels = np.empty(n, np.int64)
    o = 0
    for c, m in enumerate(counts):
        x[o:o + m] = centers[c] + rng.standard_normal((m, dim))
        labels[o:o + m] = c
        o += m
    if n_bg:
        lo, hi = centers.min(axis=0), centers.max(axis=0)
        x[o:] = lo + (hi - lo) * rng.random((n_bg, dim))
        labels[o:] = -1
    perm = rng.permutation(n)
    return x[perm], labels[perm]


# ----------------------------------------------------------------- kNN
def knn_exact(x, k, device=None, chunk=None):
    """Exact k-NN (self excluded).  Returns (indices int32 [n,k], dists float32 [n,k])."""
    import torch

    n = x.shape[0]
    if device is None:
        device = "cuda" if torch.cuda.is_available() else "cpu"
    xt = torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(device)
    sq = (xt * xt).sum(dim=1)
    if chunk is None:
        budget = 16e9 if str(device).startswith("cuda") else 1e9  # bytes of one distance tile
        chunk = max(256, min(8192, int(budget // (4 * max(n, 1)))))
    idx_out = np.empty((n, k), np.int32)
    dst_out = np.empty((n, k), np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        d2 = sq[s:e, None] - 2.0 * (xt[s:e] @ xt.T) + sq[None, :]
        d2[torch.arange(e - s, device=device), torch.arange(s, e, device=device)] = float("inf")
        vals, idx = torch.topk(d2, k, dim=1, largest=False, sorted=True)
        # exact distances for the selected pairs (avoid cancellation in the expansion)
        diff = xt[s:e, None, :] - xt[idx]
        dd = torch.sqrt((diff * diff).sum(dim=2))
        dd, order = torch.sort(dd, dim=1)
        idx = torch.gather(idx, 1, order)
        idx_out[s:e] = idx.to(torch.int32).cpu().numpy()
        dst_out[s:e] = dd.to(torch.float32).cpu().numpy()
    return idx_out, dst_out


# ----------------------------------------------------------------- weights
def smooth_knn(dists, k, n_iter=64, tol=1e-5):
    """rho = nearest distance; sigma by bisection so that sum_j exp(-(d-rho)/sigma) = log2(k)."""
    d = np.asarray(dists, np.float32)
    n = d.shape[0]
    target = np.float32(math.log2(k))
    rho = d[:, 0].copy()
    lo = np.zeros(n, np.float32)
    hi = np.full(n, np.inf, np.float32)
    mid = np.ones(n, np.float32)
    done = np.zeros(n, bool)
    gap = np.maximum(d - rho[:, None], 0.0)
    for _ in range(n_iter):
        psum = np.exp(-gap / mid[:, None]).sum(axis=1, dtype=np.float32)
        hit = np.abs(psum - target) < tol
        done |= hit
        upd = ~done
        over = upd & (psum > target)
        under = upd & ~(psum > target)
        hi = np.where(over, mid, hi)
        lo = np.where(under, mid, lo)
        mid = np.where(over, (lo + hi) / 2,
                       np.where(under, np.where(np.isinf(hi), mid * 2, (lo + hi) / 2), mid))
        if done.all():
            break
    mean_d = d.mean(axis=1)
    sigma = np.maximum(mid, np.float32(1e-3) * mean_d).astype(np.float32)
    return sigma, rho


def fuzzy_union(knn_idx, knn_dst, sigma, rho):
    """Symmetric fuzzy union W = w + w^T - w o w^T as sorted float32 CSR."""
    import scipy.sparse as sp

    n, k = knn_idx.shape
    gap = np.maximum(knn_dst - rho[:, None], 0.0).astype(np.float32)
    w = np.exp(-gap / sigma[:, None]).astype(np.float32)
    rows = np.repeat(np.arange(n, dtype=np.int64), k)
    a = sp.csr_matrix((w.ravel(), (rows, knn_idx.ravel().astype(np.int64))), shape=(n, n),
                      dtype=np.float32)
    a.sum_duplicates()
    at = a.T.tocsr()
    s = (a + at - a.multiply(at)).tocsr()
    s.sum_duplicates()
    s.eliminate_zeros()
    s.sort_indices()
    return (s.indptr.astype(np.int64), s.indices.astype(np.int32), s.data.astype(np.float32))


def fuzzy_knn_graph(x, k=15, device=None):
    idx, dst = knn_exact(x, k, device=device)
    sigma, rho = smooth_knn(dst, k)
    return fuzzy_union(idx, dst, sigma, rho)


def n_connected_components(indptr, indices):
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg

    n = indptr.size - 1
    a = sp.csr_matrix((np.ones(indices.size, np.int8), indices, indptr), shape=(n, n))
    return cg.connected_components(a, directed=False)[0]


def band_graph(n, per_row, half_band, seed):
    """Symmetric random band graph: per_row random neighbours within +-half_band."""
    import scipy.sparse as sp

    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n, dtype=np.int64), per_row)
    off = rng.integers(1, half_band + 1, rows.size) * rng.choice([-1, 1], rows.size)
    cols = np.clip(rows + off, 0, n - 1)
    keep = cols != rows
    w = rng.uniform(0.05, 1.0, keep.sum()).astype(np.float32)
    a = sp.csr_matrix((w, (rows[keep], cols[keep])), shape=(n, n))
    a.sum_duplicates()
    a = a.maximum(a.T).tocsr()
    a.sort_indices()
    return a.indptr.astype(np.int64), a.indices.astype(np.int32), a.data.astype(np.float32)


def largest_connected(indptr, indices, data):
    """Restrict a CSR graph to its largest connected component (node order kept)."""
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg

    n = indptr.size - 1
    a = sp.csr_matrix((data, indices, indptr), shape=(n, n))
    nc, lab = cg.connected_components(a, directed=False)
    if nc == 1:
        return indptr, indices, data, np.arange(n)
    keep = np.flatnonzero(lab == np.bincount(lab).argmax())
    sub = a[keep][:, keep].tocsr()
    sub.sort_indices()
    return (sub.indptr.astype(np.int64), sub.indices.astype(np.int32),
            sub.data.astype(np.float32), keep)


def components_graph(n_comp, lo, hi, dim=30, seed=0, device=None, clusters=3,
                     center_scale=0.8, spread=10.0, tiny=()):
    """Block-diagonal graph of independent components: component c has a size
    drawn log-uniformly from [lo, hi], Gaussian-mixture points in `dim`-d
    (`clusters` clusters, centres ~ N(0, center_scale^2 I)) around an offset
    ~ N(0, spread^2 I), and its own exact 15-NN fuzzy graph, restricted to its
    largest connected component.  `tiny` adds components of the given sizes
    (2..15 nodes) as small cliques.  Returns the CSR triple, the points
    (n x dim, for the meta-embedding) and the component sizes."""
    rng = np.random.default_rng(20_000 + seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_comp)).astype(np.int64)
    ips, ixs, ws, pts, out_sizes = [], [], [], [], []
    off = 0    # node offset
    noff = 0   # stored-entry offset
    for c, sz in enumerate(sizes):
        x, _ = gmm_points(int(sz), dim, np.ones(clusters), seed * 1000 + c, center_scale=center_scale)
        x = x + rng.standard_normal(dim).astype(np.float32) * spread
        ip, ix, w = fuzzy_knn_graph(x, 15, device=device)
        ip, ix, w, keep = largest_connected(ip, ix, w)
        ips.append(ip[1:] + noff if ips else ip + noff)
        ixs.append(ix + off)
        ws.append(w)
        pts.append(x[keep])
        off += keep.size
        noff += int(ip[-1])
        out_sizes.append(int(keep.size))
    for t in tiny:
        t = int(t)
        x = rng.standard_normal((t, dim)).astype(np.float32) * 0.1 + rng.standard_normal(dim).astype(np.float32) * spread
        rows, cols = np.nonzero(~np.eye(t, dtype=bool))
        ip = np.zeros(t + 1, np.int64)
        ip[1:] = np.cumsum(np.bincount(rows, minlength=t))
        ips.append(ip[1:] + noff if ips else ip + noff)
        ixs.append((cols + off).astype(np.int32))
        ws.append(np.full(rows.size, 0.5, np.float32))
        pts.append(x)
        off += t
        noff += int(rows.size)
        out_sizes.append(t)
    indptr = np.concatenate(ips).astype(np.int64)
    indices = np.concatenate(ixs).astype(np.int32)
    data = np.concatenate(ws).astype(np.float32)
    return indptr, indices, data, np.concatenate(pts).astype(np.float32), out_sizes


# ----------------------------------------------------------------- workloads
def workload_points(name, seed=0, n=None):
    """The points (float32 [n, d]) and cluster labels of the point-cloud
    workloads blobs2k, gmm100k and gmm1m, as workload() draws them."""
    rng = np.random.default_rng(10_000 + seed)
    if name == "blobs2k":
        return gmm_points(n or 2000, 50, np.ones(10), seed, center_scale=CENTER_SCALE["blobs2k"])
    if name == "gmm100k":
        sizes = rng.dirichlet(np.full(20, 2.0))
        return gmm_points(n or 100_000, 50, sizes, seed, center_scale=CENTER_SCALE["gmm100k"])
    if name == "gmm1m":
        sizes = 1.0 / np.arange(1, 41) ** 1.2
        return gmm_points(n or 1_000_000, 30, sizes, seed, center_scale=CENTER_SCALE["gmm1m"],
                          background_frac=0.05)
    raise ValueError(name)


def workload(name, seed=0, device=None, n=None):
    """Seeded inputs of the named workloads.

    blobs2k   : n=2000, 10 clusters in 50-d, equal sizes (CPU-runnable case)
    gmm100k   : n=100 000, 20 clusters in 50-d, Dirichlet(2) sizes
    gmm1m     : n=1 000 000, 40 clusters in 30-d, Zipf(1.2) sizes, 5 % background
    comp64    : 64 independent 30-d GMM 15-NN components, sizes log-uniform in [50, 200 000]
    n overrides the node count (same construction, for reduced-size runs).
    """
    if name in ("blobs2k", "gmm100k", "gmm1m"):
        x, lab = workload_points(name, seed, n)
        n = x.shape[0]
    elif name == "comp64":
        # 64 disconnected components, sizes log-uniform in [50, 200 000] (~2M nodes)
        # component offsets ~N(0, 0.5^2 I): centroid affinities exp(-d^2) stay well
        # inside the FP64 range, so the spectral meta-embedding is well posed
        indptr, indices, data, x, sizes = components_graph(64, 50, 200_000, dim=30, seed=seed,
                                                           device=device, spread=0.5)
        return dict(name=name, n=indptr.size - 1, x=x, labels=None, indptr=indptr,
                    indices=indices, data=data, sizes=sizes)
    elif name == "band1m":
        # SpMM locality reference (not a solver workload): symmetric random band
        # graph, 12 neighbours per row within +-512 rows (~24 stored per row)
        n = n or 1_000_000
        indptr, indices, data = band_graph(n, 12, 512, seed)
        return dict(name=name, n=n, x=None, labels=None, indptr=indptr, indices=indices,
                    data=data)
    else:
        raise ValueError(name)
    indptr, indices, data = fuzzy_knn_graph(x, 15, device=device)
    return dict(name=name, n=n, x=x, labels=lab, indptr=indptr, indices=indices, data=data)


# Center spread per workload.  With N(0, 10^2 I) centers and unit-variance
# clusters in 30-50 dimensions the clusters are farther apart than any k-NN
# edge reaches, so the graph splits into one component per cluster; the
# spreads below keep clear cluster structure while the graph stays connected.
CENTER_SCALE = {"blobs2k": 0.6, "gmm100k": 0.6, "gmm1m": 0.8}
