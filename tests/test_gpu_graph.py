"""GPU parity of the fuzzy k-NN graph builder (include/spectral_mi355x_graph.h)
against the reference fixtures step1_knn_exact .. step5a_pruned_exact and the
CPU oracle (oracle/oracle_graph.c), all through the C ABI.

Bars: k-NN indices bit-exact (index work); distances, sigma, rho, membership
values and W bit-exact against the oracle, which computes the same float64 /
float32 operations in the same order (the device's float32 exp restates
glibc's expf, the oracle calls it).  Against the fixtures the oracle's own
tolerances apply (sigma within 1e-5 where the fixture's bisection stopped one
step apart; near_dupes_100 step 1 is rounding noise in the reference).
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
import spectral_mi355x as si

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATASETS = [
    "blobs_50", "blobs_500", "blobs_5000", "moons_200", "circles_300",
    "near_dupes_100", "disconnected_200", "blobs_connected_200", "blobs_connected_2000",
]


def load(name):
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "graph", f"{name}.npz"),
                        allow_pickle=False))


def step3_dense(z):
    n, k = z["knn_indices"].shape
    a = sp.csr_matrix((z["s3_data"], (z["s3_row"], z["s3_col"])), shape=(n, n))
    return np.asarray(a[np.repeat(np.arange(n), k), z["knn_indices"].ravel()],
                      np.float32).reshape(n, k)


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    c = si.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", DATASETS)
def test_knn_matches_oracle_and_fixture(ctx, name):
    print(f"case knn {name}", flush=True)
    z = load(name)
    k = int(z["n_neighbors"])
    idx, dist = si.knn_exact(ctx, z["x"], k)
    oi, od = O.knn_exact(z["x"], k)
    np.testing.assert_array_equal(idx, oi)
    np.testing.assert_allclose(dist, od, rtol=1.2e-7, atol=0)
    if name != "near_dupes_100":
        np.testing.assert_array_equal(idx, z["knn_indices"])
        np.testing.assert_allclose(dist, z["knn_dists"], rtol=1.2e-7, atol=0)
    else:  # the 90 ordinary points
        np.testing.assert_array_equal(idx[:90], z["knn_indices"][:90])


@pytest.mark.parametrize("d,k", [(3, 5), (17, 15), (30, 15), (50, 15), (64, 32), (9, 64)])
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_knn_random_shapes_match_oracle(ctx, d, k, precision):
    print(f"case knn random d={d} k={k} {precision}", flush=True)
    rng = np.random.default_rng(d * 100 + k)
    n = 2500
    x = rng.standard_normal((n, d))
    x[7] = x[1234]                    # exact duplicate: tie at distance 0
    x[100:110] = x[100]               # ten copies: ties broken by index
    if precision == "f32":
        x = x.astype(np.float32)
    idx, dist = si.knn_exact(ctx, x, k)
    oi, od = O.knn_exact(x.astype(np.float64), k)
    np.testing.assert_array_equal(idx, oi)
    np.testing.assert_allclose(dist, od, rtol=1.2e-7, atol=0)
    assert idx[7, 0] == 1234 and idx[1234, 0] == 7
    m = min(k, 9)
    np.testing.assert_array_equal(idx[100, :m], np.arange(101, 101 + m))


@pytest.mark.parametrize("name", DATASETS)
def test_smooth_knn_dist_matches_oracle_and_fixture(ctx, name):
    print(f"case smooth_knn {name}", flush=True)
    z = load(name)
    sig, rho = si.smooth_knn_dist(ctx, z["knn_dists"])
    osig, orho = O.smooth_knn_dist(z["knn_dists"])
    np.testing.assert_array_equal(rho, orho)
    np.testing.assert_array_equal(sig, osig)
    np.testing.assert_array_equal(rho, z["rhos"])
    np.testing.assert_array_equal(np.isinf(sig), np.isinf(z["sigmas"]))
    fin = np.isfinite(z["sigmas"])
    np.testing.assert_allclose(sig[fin], z["sigmas"][fin], rtol=1e-5)


@pytest.mark.parametrize("name", DATASETS)
def test_membership_matches_fixture(ctx, name):
    print(f"case membership {name}", flush=True)
    z = load(name)
    v = si.membership_strengths(ctx, z["knn_indices"], z["knn_dists"], z["sigmas"], z["rhos"])
    np.testing.assert_array_equal(v, step3_dense(z))


@pytest.mark.parametrize("name", DATASETS)
def test_fuzzy_union_and_prune_match_fixture(ctx, name):
    print(f"case fuzzy_union {name}", flush=True)
    z = load(name)
    v3 = step3_dense(z)
    n = z["x"].shape[0]
    ip, ix, w = si.fuzzy_union(ctx, z["knn_indices"], v3, n_epochs=0).to_host()
    np.testing.assert_array_equal(ip, z["s4_indptr"])
    np.testing.assert_array_equal(ix, z["s4_indices"])
    np.testing.assert_array_equal(w, z["s4_data"])
    ip, ix, w = si.fuzzy_union(ctx, z["knn_indices"], v3, n_epochs=-1).to_host()
    np.testing.assert_array_equal(ip, z["s5_indptr"])
    np.testing.assert_array_equal(ix, z["s5_indices"])
    np.testing.assert_array_equal(w, z["s5_data"])


@pytest.mark.parametrize("name", DATASETS)
def test_fuzzy_graph_from_points(ctx, name):
    """Raw points -> W in one device call: bit-exact against the oracle's
    chain; against step5a_pruned_exact the structure exactly and the values to
    the sigma tolerance (near_dupes_100: oracle only)."""
    print(f"case fuzzy_graph {name}", flush=True)
    z = load(name)
    k = int(z["n_neighbors"])
    ip, ix, w = si.fuzzy_graph(ctx, z["x"], k).to_host()
    oip, oix, ow = O.fuzzy_graph(z["x"], k)
    np.testing.assert_array_equal(ip, oip)
    np.testing.assert_array_equal(ix, oix)
    np.testing.assert_array_equal(w, ow)
    if name != "near_dupes_100":
        np.testing.assert_array_equal(ip, z["s5_indptr"])
        np.testing.assert_array_equal(ix, z["s5_indices"])
        np.testing.assert_allclose(w, z["s5_data"], rtol=2e-5, atol=0)


def test_fuzzy_graph_feeds_the_eigensolver(ctx):
    """The device-built W of blobs_connected_2000 gives the spectrum of the
    fixture's step5a_pruned_exact W (numpy eigh of its dense Laplacian; 1e-4
    relative, W agrees to 2e-5)."""
    z = load("blobs_connected_2000")
    ip, ix, w = si.fuzzy_graph(ctx, z["x"], 15).to_host()
    lap = si.Laplacian(ctx, ip, ix, w)
    ev, vec, rep = si.solve_eigenproblem(lap, 2, 42, si.SpectralInitConfig(tol=1e-6))
    lip, lix, ldat = O.laplacian_from_graph(z["s5_indptr"], z["s5_indices"], z["s5_data"])[0]
    n = lip.size - 1
    dense = sp.csr_matrix((ldat, lix, lip), shape=(n, n)).toarray()
    ev_ref = np.linalg.eigvalsh(dense)[:3]
    assert abs(ev[0]) < 1e-9
    np.testing.assert_allclose(ev[1:], ev_ref[1:], rtol=1e-4)
    assert rep.max_residual <= 1e-6


def test_fuzzy_graph_100k_properties_and_sampled_rows(ctx):
    """n = 100 000 points in 50-d (20 Gaussian clusters): the sampled rows'
    neighbour lists equal a float64 brute force; steps 2-5 from the device's
    k-NN lists equal the oracle's bit for bit; W is symmetric with weights in
    (0, 1]."""
    print("case fuzzy_graph 100k", flush=True)
    rng = np.random.default_rng(11)
    n, d, k = 100_000, 50, 15
    centres = rng.standard_normal((20, d)) * 4.0
    x = centres[rng.integers(0, 20, n)] + rng.standard_normal((n, d))
    idx, dist = si.knn_exact(ctx, x, k)
    rows = rng.choice(n, 64, replace=False)
    for i in rows:
        d2 = ((x - x[i]) ** 2).sum(1)
        d2[i] = np.inf
        ref = np.argsort(d2, kind="stable")[:k]
        np.testing.assert_array_equal(idx[i], ref)
    sig, rho = si.smooth_knn_dist(ctx, dist)
    osig, orho = O.smooth_knn_dist(dist)
    np.testing.assert_array_equal(sig, osig)
    np.testing.assert_array_equal(rho, orho)
    v = si.membership_strengths(ctx, idx, dist, sig, rho)
    np.testing.assert_array_equal(v, O.membership(idx, dist, osig, orho))
    g = si.fuzzy_union(ctx, idx, v, n_epochs=-1).to_host()
    og = O.fuzzy_union(idx, v, 200)
    for a, b in zip(g, og):
        np.testing.assert_array_equal(a, b)
    ip, ix, w = g
    W = sp.csr_matrix((w, ix, ip), shape=(n, n))
    assert (W != W.T).nnz == 0
    assert w.min() > 0 and w.max() <= 1.0
    assert np.all(np.diff(ip) >= 1)
    full = si.fuzzy_graph(ctx, x, k).to_host()
    for a, b in zip(full, og):
        np.testing.assert_array_equal(a, b)


def test_fuzzy_graph_1m_points_properties(ctx):
    """The 1M-point, 30-d workload at full size: sampled neighbour lists equal a
    float64 brute force, sigma/rho equal the oracle's on the device's lists, W
    is bitwise symmetric with weights in (0, 1] and every row non-empty."""
    import graphs
    print("case fuzzy_graph 1m", flush=True)
    x, _ = graphs.workload_points("gmm1m", 0)
    n, k = x.shape[0], 15
    idx, dist = si.knn_exact(ctx, x, k)
    x64 = x.astype(np.float64)
    for i in np.random.default_rng(2).choice(n, 16, replace=False):
        d2 = ((x64 - x64[i]) ** 2).sum(1)
        d2[i] = np.inf
        np.testing.assert_array_equal(idx[i], np.argsort(d2, kind="stable")[:k])
    sig, rho = si.smooth_knn_dist(ctx, dist)
    osig, orho = O.smooth_knn_dist(dist)
    np.testing.assert_array_equal(sig, osig)
    np.testing.assert_array_equal(rho, orho)
    ip, ix, w = si.fuzzy_graph(ctx, x, k).to_host()
    W = sp.csr_matrix((w, ix, ip), shape=(n, n))
    assert (W != W.T).nnz == 0
    assert w.min() > 0 and w.max() <= 1.0
    assert np.all(np.diff(ip) >= 1)
    # columns strictly ascending inside every row
    within = np.ones(ix.size - 1, bool)
    within[ip[1:-1] - 1] = False
    assert np.all(np.diff(ix.astype(np.int64))[within] > 0)


def test_fuzzy_union_edge_cases(ctx):
    """All memberships zero: a graph without edges; k = 1; missing (-1) entries."""
    n, k = 50, 4
    idx = np.stack([(np.arange(k) + i + 1) % n for i in range(n)]).astype(np.int32)
    ip, ix, w = si.fuzzy_union(ctx, idx, np.zeros((n, k), np.float32), n_epochs=0).to_host()
    assert np.all(ip == 0) and ix.size == 0 and w.size == 0
    v = np.random.default_rng(1).uniform(0.1, 1.0, (n, k)).astype(np.float32)
    idx[::5, 2] = -1
    for kk in (1, k):
        g = si.fuzzy_union(ctx, idx[:, :kk], v[:, :kk], n_epochs=0).to_host()
        og = O.fuzzy_union(idx[:, :kk], v[:, :kk], 0)
        for a, b in zip(g, og):
            np.testing.assert_array_equal(a, b)


def test_graph_builder_errors(ctx):
    x = np.zeros((10, 65), np.float32)
    with pytest.raises(si.Unsupported):
        si.knn_exact(ctx, x, 3)
    with pytest.raises(ValueError):
        si.knn_exact(ctx, np.zeros((5, 2)), 5)
    with pytest.raises(si.Unsupported):
        si.knn_exact(ctx, np.zeros((100, 2)), 65)
