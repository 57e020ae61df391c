"""The locality order and the component labels computed on the device.

rcm_order_device must return the host order of reorder.cpp bit for bit (the
host order is pinned by tests/test_reorder.py against a sequential Python
restatement of Cuthill-McKee); components_device must return the labels of
the oracle's find_components (src/components.rs:7-44: BFS in node order, so
components are numbered by their smallest node).  Graphs: random graphs with
unsorted rows and repeated columns, several components with isolated nodes
and self loops, hubs whose new-node groups exceed one wavefront (the
segmented sort of the last sweep), a star, a complete graph, a ring, many
small components with shuffled ids, and the fixture graphs.

Connected graphs of at least 65536 nodes and 8 entries per node are stored in
the clustering order
(k-means clusters of a random-walk embedding): a permutation, the same on
every build, undone by every entry point -- the exported Laplacian, the SpMM
and the eigenpairs are those of the input order -- while disconnected graphs,
mesh-like graphs and the explicit RCM flag keep reverse Cuthill-McKee.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from conftest import load_fixture
from test_reorder import random_graph

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
si = pytest.importorskip("spectral_mi355x")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = si.Context(0)
    yield c
    c.close()


def _csr(a):
    a = a.tocsr()
    a.sort_indices()
    return a.indptr.astype(np.int64), a.indices.astype(np.int32)


def _sym(rows, cols, n):
    a = sp.coo_matrix((np.ones(rows.size), (rows, cols)), shape=(n, n)).tocsr()
    return _csr(a + a.T)


def multi_component(seed=3):
    # three random components, isolated nodes, self loops, ids shuffled
    rng = np.random.default_rng(seed)
    parts, base = [], 0
    for sz, k in ((4000, 3), (700, 2), (90, 4)):
        ip, ix = random_graph(sz, k, int(rng.integers(1 << 30)))
        r = np.repeat(np.arange(sz), np.diff(ip))
        parts.append((r + base, ix.astype(np.int64) + base))
        base += sz
    n = base + 25
    rows = np.concatenate([p[0] for p in parts] + [np.arange(0, 400, 7)])
    cols = np.concatenate([p[1] for p in parts] + [np.arange(0, 400, 7)])
    relabel = rng.permutation(n)
    return _sym(relabel[rows], relabel[cols], n)


def hubs(n=6000, seed=5):
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n), 3)
    cols = rng.integers(0, n, 3 * n)
    extra_r, extra_c = [], []
    for h in range(6):
        nb = rng.choice(n, 300 + 400 * h, replace=False)
        extra_r.append(np.full(nb.size, h * 997 % n))
        extra_c.append(nb)
    rows = np.concatenate([rows] + extra_r)
    cols = np.concatenate([cols] + extra_c)
    keep = rows != cols
    return _sym(rows[keep], cols[keep], n)


def star(n=5000):
    return _sym(np.zeros(n - 1, np.int64), np.arange(1, n), n)


def complete(n=150):
    r, c = np.nonzero(~np.eye(n, dtype=bool))
    return _csr(sp.coo_matrix((np.ones(r.size), (r, c)), shape=(n, n)))


def ring(n=10000):
    i = np.arange(n)
    return _sym(i, (i + 1) % n, n)


def many_rings(seed=9):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(3, 40, 400)
    rows, cols, base = [], [], 0
    for s in sizes:
        i = np.arange(s)
        rows.append(base + i)
        cols.append(base + (i + 1) % s)
        base += s
    relabel = rng.permutation(base)
    return _sym(relabel[np.concatenate(rows)], relabel[np.concatenate(cols)], base)


def fixture(name):
    z = load_fixture(name)
    return z["w_indptr"].astype(np.int64), z["w_indices"].astype(np.int32)


GRAPHS = {
    "random_dup": lambda: random_graph(5000, 3, 7, True),
    "random_large": lambda: random_graph(120_000, 4, 11, False),
    "multi_component": multi_component,
    "hubs": hubs,
    "star": star,
    "complete150": complete,
    "ring10000": ring,
    "many_rings": many_rings,
    "blobs_5000": lambda: fixture("blobs_5000"),
    "disconnected_200": lambda: fixture("disconnected_200"),
    "blobs_connected_2000": lambda: fixture("blobs_connected_2000"),
}


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_device_components_match_find_components(ctx, name):
    print(f"case components {name}", flush=True)
    ip, ix = GRAPHS[name]()
    lab, nc = si.components_device(ctx, ip, ix)
    ref, rnc = O.find_components(ip, ix)
    assert nc == rnc
    np.testing.assert_array_equal(lab, ref)


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_device_rcm_is_the_host_order(ctx, name):
    print(f"case rcm {name}", flush=True)
    ip, ix = GRAPHS[name]()
    perm = si.rcm_order_device(ctx, ip, ix)
    np.testing.assert_array_equal(perm, si.rcm_order(ip, ix))


def test_device_rcm_repeated_calls_agree(ctx):
    ip, ix = multi_component(4)
    p1 = si.rcm_order_device(ctx, ip, ix)
    p2 = si.rcm_order_device(ctx, ip, ix)
    np.testing.assert_array_equal(p1, p2)


def _storage_matches(ctx, lap, ip, ix, w):
    """The stored CSR is P W P^T of the input and its FP32 values are the input
    order's, carried along bit for bit."""
    n = ip.size - 1
    rp, col, a32, perm, order = si.laplacian_storage(lap)
    assert np.array_equal(np.sort(perm), np.arange(n))
    plain = si.Laplacian(ctx, ip, ix, w, reorder=False)
    _, _, pa32, _, _ = si.laplacian_storage(plain)
    a = sp.csr_matrix((pa32, ix, ip), shape=(n, n))[perm][:, perm].tocsr()
    a.sort_indices()
    np.testing.assert_array_equal(rp, a.indptr)
    np.testing.assert_array_equal(col, a.indices)
    np.testing.assert_array_equal(a32, a.data)
    return perm, order


def test_permuted_rows_of_every_length(ctx):
    # rows of up to 64 entries are ranked by one register per lane, rows of
    # 65..256 by four, longer rows by the segmented sort of just those rows
    print("case permuted csr row lengths", flush=True)
    n = 6000
    ip, ix = random_graph(n, 4, 21, False)
    rows = [np.repeat(np.arange(n), np.diff(ip))]
    cols = [ix.astype(np.int64)]
    rng = np.random.default_rng(22)
    for h, deg in enumerate((63, 64, 65, 128, 200, 255, 256, 257, 700)):
        nb = rng.choice(np.arange(100, n), deg, replace=False)
        rows.append(np.full(deg, h * 11))
        cols.append(nb)
    r, c = np.concatenate(rows), np.concatenate(cols)
    keep = r != c
    ip, ix = _sym(r[keep], c[keep], n)
    w = rng.uniform(0.1, 1.0, ix.size).astype(np.float32)
    a = sp.csr_matrix((w, ix, ip), shape=(n, n))
    a = ((a + a.T) * 0.5).tocsr()
    a.sort_indices()
    ip, ix, w = a.indptr.astype(np.int64), a.indices.astype(np.int32), a.data.astype(np.float32)
    lens = np.diff(ip)
    assert lens.max() > 256 and ((lens > 64) & (lens <= 256)).sum() >= 5
    lap = si.Laplacian(ctx, ip, ix, w)
    perm, order = _storage_matches(ctx, lap, ip, ix, w)
    assert order == si.ORDER_RCM


def test_cluster_order_for_large_connected_graphs(ctx):
    print("case cluster order random_large", flush=True)
    ip, ix = random_graph(120_000, 6, 11, False)
    w = np.random.default_rng(3).uniform(0.1, 1.0, ix.size).astype(np.float32)
    a = sp.csr_matrix((w, ix, ip))
    a = ((a + a.T) * 0.5).tocsr()
    a.sort_indices()
    ip, ix, w = a.indptr.astype(np.int64), a.indices.astype(np.int32), a.data.astype(np.float32)
    lap = si.Laplacian(ctx, ip, ix, w)
    perm, order = _storage_matches(ctx, lap, ip, ix, w)
    assert order == si.ORDER_CLUSTER
    assert not np.array_equal(perm, si.rcm_order(ip, ix))
    # deterministic: a second build stores the same order
    perm2, _ = _storage_matches(ctx, si.Laplacian(ctx, ip, ix, w), ip, ix, w)
    np.testing.assert_array_equal(perm, perm2)
    # the explicit RCM flag keeps reverse Cuthill-McKee
    lap_r = si.Laplacian(ctx, ip, ix, w, reorder="rcm")
    perm_r, order_r = _storage_matches(ctx, lap_r, ip, ix, w)
    assert order_r == si.ORDER_RCM
    np.testing.assert_array_equal(perm_r, si.rcm_order(ip, ix))
    # the order is invisible: the exported L and the SpMM are the input order's
    deg, sq, (lip, lix, ld) = lap.export()
    (oip, oix, od), odeg, _ = O.laplacian_from_graph(ip, ix, w)
    np.testing.assert_array_equal(lip, oip)
    np.testing.assert_array_equal(lix, oix)
    np.testing.assert_array_equal(ld, od)
    x = np.random.default_rng(5).standard_normal((ip.size - 1, 3))
    xd = torch.from_numpy(x).cuda()
    y1, y2 = torch.empty_like(xd), torch.empty_like(xd)
    lap.spmm(xd, y1, si.SI_F64)
    lap_r.spmm(xd, y2, si.SI_F64)
    np.testing.assert_allclose(y1.cpu().numpy(), y2.cpu().numpy(), rtol=1e-13, atol=1e-13)
    # and the eigenpairs agree with the RCM-stored operator's
    ev1, v1, r1 = si.solve_eigenproblem(lap, 2, 42, si.SpectralInitConfig(tol=1e-8))
    ev2, v2, r2 = si.solve_eigenproblem(lap_r, 2, 42, si.SpectralInitConfig(tol=1e-8))
    assert r1.max_residual <= 1e-8 and r2.max_residual <= 1e-8
    np.testing.assert_allclose(ev1, ev2, rtol=1e-9, atol=1e-12)


def test_disconnected_large_graphs_keep_rcm(ctx):
    print("case cluster order not for several components", flush=True)
    ip, ix = random_graph(70_000, 4, 12, False)
    ip2, ix2 = random_graph(5000, 3, 13, False)
    ip = np.concatenate([ip, ip[-1] + ip2[1:]])
    ix = np.concatenate([ix, ix2 + 70_000]).astype(np.int32)
    w = np.ones(ix.size, np.float32)
    lap = si.Laplacian(ctx, ip, ix, w)
    perm, order = _storage_matches(ctx, lap, ip, ix, w)
    assert order == si.ORDER_RCM
    np.testing.assert_array_equal(perm, si.rcm_order(ip, ix))


def test_mesh_like_large_graphs_keep_rcm(ctx):
    # average degree 2: RCM's band (2 for a ring) serves the SpMM and level 2
    print("case cluster order not for a ring", flush=True)
    ip, ix = ring(70_000)
    w = np.ones(ix.size, np.float32)
    lap = si.Laplacian(ctx, ip, ix, w)
    assert si.laplacian_order(lap) == si.ORDER_RCM
