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
