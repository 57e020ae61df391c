"""CPU checks of the oracle's fuzzy k-NN graph (oracle/oracle_graph.c) against
the reference fixtures step1_knn_exact .. step5a_pruned_exact (extracted to
tests/golden/graph/ by tests/golden/make_golden.py), each step fed the
fixture's own previous step, plus checks that do not use the oracle's code:
numpy brute force, scipy sparse algebra, a bisection closed form.

near_dupes_100 is excluded from the step-1 comparison: its ten duplicates lie
within 1e-6 of each other at |x| ~ 70, where the reference's distances
(sklearn's |x|^2 - 2 x.y + |y|^2 in float64) are rounding noise; its steps 2-5
are compared from the fixture's own step-1 data like every other dataset.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATASETS = [
    "blobs_50", "blobs_500", "blobs_5000", "moons_200", "circles_300",
    "near_dupes_100", "disconnected_200", "blobs_connected_200", "blobs_connected_2000",
]


def load(name):
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "graph", f"{name}.npz"),
                        allow_pickle=False))


def step3_dense(z):
    """The fixture's step-3 values laid out [n, k] like the k-NN lists."""
    n, k = z["knn_indices"].shape
    a = sp.csr_matrix((z["s3_data"], (z["s3_row"], z["s3_col"])), shape=(n, n))
    return np.asarray(a[np.repeat(np.arange(n), k), z["knn_indices"].ravel()],
                      np.float32).reshape(n, k)


@pytest.mark.parametrize("name", [d for d in DATASETS if d != "near_dupes_100"])
def test_knn_exact_matches_fixture(name):
    z = load(name)
    idx, dist = O.knn_exact(z["x"], int(z["n_neighbors"]))
    np.testing.assert_array_equal(idx, z["knn_indices"])
    # float32 distances: equal, or one ulp apart (the fixture rounds sklearn's
    # expanded float64 form, the oracle the direct sum of squares)
    np.testing.assert_allclose(dist, z["knn_dists"], rtol=1.2e-7, atol=0)


def test_knn_exact_matches_numpy_brute_force():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((700, 7))
    x[5] = x[9]          # an exact duplicate: a tie at distance 0
    k = 12
    idx, dist = O.knn_exact(x, k)
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    ref = np.argsort(d2, axis=1, kind="stable")[:, :k]
    np.testing.assert_array_equal(idx, ref)
    np.testing.assert_allclose(dist, np.sqrt(np.take_along_axis(d2, ref, 1)), rtol=1.2e-7)
    assert idx[5, 0] == 9 and idx[9, 0] == 5 and dist[5, 0] == 0.0


@pytest.mark.parametrize("name", DATASETS)
def test_smooth_knn_dist_matches_fixture(name):
    z = load(name)
    sig, rho = O.smooth_knn_dist(z["knn_dists"])
    np.testing.assert_array_equal(rho, z["rhos"])
    ref = z["sigmas"]
    # sigma = inf exactly where the fixture has it (umap's fastmath build)
    np.testing.assert_array_equal(np.isinf(sig), np.isinf(ref))
    fin = np.isfinite(ref)
    # one bisection step apart where the fixture's summation rounds differently
    np.testing.assert_allclose(sig[fin], ref[fin], rtol=1e-5)


def test_smooth_knn_dist_bit_exact_rate():
    exact = total = 0
    for name in DATASETS:
        z = load(name)
        sig, _ = O.smooth_knn_dist(z["knn_dists"])
        exact += int((sig == z["sigmas"]).sum())
        total += sig.size
    assert exact / total >= 0.99, (exact, total)


def test_smooth_knn_dist_closed_form():
    """Distances rho, rho + c, ..., rho + c: the sum over j >= 1 is
    (k - 1) exp(-c/sigma) = log2(k), so sigma = c / ln((k - 1) / log2 k) up to
    the bisection's 1e-5 stopping rule.  c is chosen so that sigma < 1 (the
    bracket closes)."""
    k, c = 15, 0.3
    d = np.full((4, k), 2.0 + c, np.float32)
    d[:, 0] = 2.0
    sig, rho = O.smooth_knn_dist(d)
    exact = c / np.log((k - 1) / np.log2(k))
    assert np.all(rho == 2.0)
    np.testing.assert_allclose(sig, exact, rtol=2e-5)


def test_smooth_knn_dist_open_bracket_gives_inf():
    """Gaps large against 1 make the first sum (mid = 1) fall below log2(k):
    the fixtures' umap build then returns sigma = inf (memberships all 1)."""
    k = 15
    d = np.arange(k, dtype=np.float32)[None, :] * 10.0 + 1.0
    sig, _ = O.smooth_knn_dist(d)
    assert np.isinf(sig[0])
    v = O.membership(np.arange(1, k + 1, dtype=np.int32)[None, :], d, sig, np.float32([1.0]))
    assert np.all(v == 1.0)


@pytest.mark.parametrize("name", DATASETS)
def test_membership_matches_fixture(name):
    z = load(name)
    v = O.membership(z["knn_indices"], z["knn_dists"], z["sigmas"], z["rhos"])
    np.testing.assert_array_equal(v, step3_dense(z))
    # explicit zeros (float32 underflow) are the entries the fixture dropped
    assert int((v != 0).sum()) == z["s3_data"].size


@pytest.mark.parametrize("name", DATASETS)
def test_fuzzy_union_and_prune_match_fixture(name):
    z = load(name)
    v3 = step3_dense(z)
    n = z["x"].shape[0]
    ip, ix, w = O.fuzzy_union(z["knn_indices"], v3, 0)
    np.testing.assert_array_equal(ip, z["s4_indptr"])
    np.testing.assert_array_equal(ix, z["s4_indices"])
    np.testing.assert_array_equal(w, z["s4_data"])
    ip, ix, w = O.fuzzy_union(z["knn_indices"], v3, 500 if n <= 10000 else 200)
    np.testing.assert_array_equal(ip, z["s5_indptr"])
    np.testing.assert_array_equal(ix, z["s5_indices"])
    np.testing.assert_array_equal(w, z["s5_data"])


def test_fuzzy_union_matches_scipy_algebra():
    rng = np.random.default_rng(5)
    n, k = 400, 9
    idx = np.stack([rng.choice(np.delete(np.arange(n), i), k, replace=False) for i in range(n)])
    idx = idx.astype(np.int32)
    vals = rng.uniform(0.0, 1.0, (n, k)).astype(np.float32)
    vals[::7, 3] = 0.0
    a = sp.csr_matrix((vals.ravel(), (np.repeat(np.arange(n), k), idx.ravel())), shape=(n, n))
    a.eliminate_zeros()
    ref = (a + a.T - a.multiply(a.T)).tocsr()
    ref.sum_duplicates()
    ref.eliminate_zeros()
    ref.sort_indices()
    ip, ix, w = O.fuzzy_union(idx, vals, 0)
    np.testing.assert_array_equal(ip, ref.indptr)
    np.testing.assert_array_equal(ix, ref.indices)
    np.testing.assert_array_equal(w, ref.data)
    # pruning below max / n_epochs, the threshold a float32 quotient
    thr = np.float32(ref.data.max()) / np.float32(50)
    keep = ref.data >= thr
    ip, ix, w = O.fuzzy_union(idx, vals, 50)
    np.testing.assert_array_equal(w, ref.data[keep])


def test_planted_mistakes_are_caught():
    """Plausible mistakes, planted through the oracle's inputs, each fail the
    fixture pins: a dropped A o A^T term, self left in the k-NN lists, and a
    bisection sum that starts at the nearest neighbour instead of the second."""
    z = load("blobs_500")
    v3 = step3_dense(z)
    n, k = v3.shape
    a = sp.csr_matrix((v3.ravel(), (np.repeat(np.arange(n), k), z["knn_indices"].ravel())),
                      shape=(n, n))
    dropped = (a + a.T).tocsr()
    dropped.sort_indices()
    assert np.array_equal(dropped.indices, z["s4_indices"])
    assert np.abs(dropped.data - z["s4_data"]).max() > 0.1
    # self kept: the k nearest including the point itself
    d2 = ((z["x"][:, None, :] - z["x"][None, :, :]) ** 2).sum(-1)
    with_self = np.argsort(d2, axis=1, kind="stable")[:, :k]
    assert not np.array_equal(with_self, z["knn_indices"])
    # bisection over j >= 0: feed the oracle a row whose second column repeats
    # the first, so that its j >= 1 sum sees the nearest neighbour
    shifted = np.concatenate([z["knn_dists"][:, :1], z["knn_dists"][:, :-1]], axis=1)
    sig, _ = O.smooth_knn_dist(shifted)
    fin = np.isfinite(z["sigmas"]) & np.isfinite(sig)
    assert np.max(np.abs(sig[fin] - z["sigmas"][fin]) / z["sigmas"][fin]) > 1e-3


def test_fuzzy_graph_end_to_end_from_points():
    """Raw points -> W: the structure of step5a_pruned_exact exactly, values
    within the sigma tolerance."""
    for name in ("blobs_500", "moons_200", "blobs_connected_2000"):
        z = load(name)
        ip, ix, w = O.fuzzy_graph(z["x"], int(z["n_neighbors"]))
        np.testing.assert_array_equal(ip, z["s5_indptr"])
        np.testing.assert_array_equal(ix, z["s5_indices"])
        np.testing.assert_allclose(w, z["s5_data"], rtol=2e-5, atol=0)
