"""GPU parity of the FP64 residual, the gate of every escalation level.

si_max_eigenpair_residual (metrics.rs:73-92 on the device) is compared value
for value with the oracle on vectors that are NOT eigenvectors (closed forms
and random blocks, 1e-12 relative), NaN propagation and scale invariance are
checked, and every solver's reported residual is recomputed independently by
the oracle from the returned pairs.  The oracle residual itself is pinned in
tests/test_oracle_residual.py.
"""
import math

import numpy as np
import pytest

import oracle as O
from conftest import load_fixture
from helpers import path_graph, random_symmetric_graph

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


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()


def test_residual_p3_closed_form(ctx):
    # normalized Laplacian of the path P3: L e1 = (1, -1/sqrt 2, 0), lambda = 0
    ip, ix, w = path_graph(3)
    lap = si.Laplacian(ctx, ip, ix, w)
    e1 = np.array([[1.0], [0.0], [0.0]])
    got = si.max_eigenpair_residual(lap, np.array([0.0]), _dev(e1))
    assert got == pytest.approx(math.sqrt(1.5), rel=1e-15)
    # lambda = 1: L e1 - e1 = (0, -1/sqrt 2, 0)
    got = si.max_eigenpair_residual(lap, np.array([1.0]), _dev(e1))
    assert got == pytest.approx(math.sqrt(0.5), rel=1e-15)
    # scale invariance
    for c in (1e-3, 7.5):
        got = si.max_eigenpair_residual(lap, np.array([0.0]), _dev(c * e1))
        assert got == pytest.approx(math.sqrt(1.5), rel=1e-14)
    # worst column, in either position (metrics.rs:456-476)
    sq = np.sqrt(np.array([1.0, 2.0, 1.0]))
    triv = (sq / np.linalg.norm(sq))[:, None]
    both = np.hstack([triv, e1])
    assert si.max_eigenpair_residual(lap, np.zeros(2), _dev(both)) == pytest.approx(math.sqrt(1.5))
    assert si.max_eigenpair_residual(lap, np.zeros(2), _dev(both[:, ::-1])) == pytest.approx(
        math.sqrt(1.5))
    assert si.max_eigenpair_residual(lap, np.zeros(1), _dev(triv)) < 1e-15


def test_residual_propagates_nan(ctx):
    ip, ix, w = path_graph(3)
    lap = si.Laplacian(ctx, ip, ix, w)
    e1 = np.array([[1.0], [0.0], [0.0]])
    nanv = np.hstack([e1, np.array([[np.nan], [0.0], [0.0]])])
    assert math.isnan(si.max_eigenpair_residual(lap, np.zeros(2), _dev(nanv)))
    assert math.isnan(si.max_eigenpair_residual(lap, np.zeros(2), _dev(nanv[:, ::-1])))
    assert math.isnan(si.max_eigenpair_residual(lap, np.array([0.0, np.nan]),
                                                _dev(np.hstack([e1, e1]))))


@pytest.mark.parametrize("n,k,reorder", [(3000, 1, False), (3000, 3, False), (6000, 5, None),
                                         (6000, 16, None)])
def test_residual_matches_oracle_on_non_eigenvectors(ctx, n, k, reorder):
    # n = 6000 takes the locality order: vectors are mapped through the permutation
    print(f"case residual n={n} k={k}", flush=True)
    ip, ix, w = random_symmetric_graph(n, 10, n + k)
    lap = si.Laplacian(ctx, ip, ix, w, reorder=reorder)
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    rng = np.random.default_rng(k)
    v = rng.standard_normal((n, k)) * rng.uniform(1e-3, 1e3, k)[None, :]
    lam = rng.uniform(0.0, 2.0, k)
    got = si.max_eigenpair_residual(lap, lam, _dev(v))
    want = O.max_eigenpair_residual(L, lam, v)
    assert got == pytest.approx(want, rel=1e-12)
    # per column: the worst column is each one in turn
    for c in range(min(k, 4)):
        g1 = si.max_eigenpair_residual(lap, lam[c:c + 1], _dev(v[:, c:c + 1]))
        assert g1 == pytest.approx(O.max_eigenpair_residual(L, lam[c:c + 1], v[:, c:c + 1]),
                                   rel=1e-12)


@pytest.fixture(scope="module")
def gmm20k():
    import graphs
    w = graphs.workload("gmm100k", seed=3, n=20000)
    L, _, _ = O.laplacian_from_graph(w["indptr"], w["indices"], w["data"])
    return w, L


def test_fast_solver_reported_residual_is_the_true_one(ctx, gmm20k):
    # the FP32-bulk solver reports ||(L Y) C - theta Y C||: recomputed by the
    # oracle as ||L v - lambda v|| / ||v|| from the returned pairs
    w, L = gmm20k
    lap = si.Laplacian(ctx, w["indptr"], w["indices"], w["data"])
    for seed in (42, 7):
        ev, vec, rep = si.fast_solve(lap, 2, seed, si.SpectralInitConfig(tol=1e-6))
        true = O.max_eigenpair_residual(L, ev, vec.cpu().numpy())
        assert rep.max_residual <= 1e-6 and true <= 1e-6
        assert rep.max_residual == pytest.approx(true, rel=1e-6)
        # and the device's own residual entry point agrees
        dev = si.max_eigenpair_residual(lap, ev, vec)
        assert dev == pytest.approx(true, rel=1e-10)


@pytest.mark.parametrize("name", ["blobs_connected_2000", "blobs_5000"])
def test_chain_and_plain_reported_residuals_are_true(ctx, name):
    z = load_fixture(name)
    ip, ix, w = z["w_indptr"], z["w_indices"], z["w_data"]
    lap = si.Laplacian(ctx, ip, ix, w)
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    ev, vec, rep = si.solve_eigenproblem(lap, 2, 42, si.SpectralInitConfig(tol=1e-6))
    true = O.max_eigenpair_residual(L, ev, vec.cpu().numpy())
    assert rep.max_residual == pytest.approx(true, rel=1e-6)
    ev, vec, rep = si.lobpcg_solve(lap, 2, 42)
    true = O.max_eigenpair_residual(L, ev, vec.cpu().numpy())
    assert rep.max_residual == pytest.approx(true, rel=1e-6)


@pytest.mark.parametrize("nc", [8, 25])
def test_solve_eigenproblem_many_components(ctx, nc):
    # k = nc + 1 columns beyond the fused update's 8 (nc = 8) and beyond the
    # FP32-bulk solver's 32-column finish (nc = 25: the chain moves on)
    print(f"case chain n_components={nc}", flush=True)
    z = load_fixture("blobs_connected_2000")
    ip, ix, w = z["w_indptr"], z["w_indices"], z["w_data"]
    lap = si.Laplacian(ctx, ip, ix, w)
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    ev, vec, rep = si.solve_eigenproblem(lap, nc, 42, si.SpectralInitConfig(tol=1e-6))
    dense = np.linalg.eigvalsh(O.spmm(L, np.eye(2000)))
    k = nc + 1
    assert ev.shape == (k,)
    np.testing.assert_allclose(ev, dense[:k], atol=1e-6)
    true = O.max_eigenpair_residual(L, ev, vec.cpu().numpy())
    assert true <= 1e-6 and rep.max_residual == pytest.approx(true, rel=1e-6)
    if nc == 8:
        assert rep.level == 1


def test_spectral_init_eight_components_disconnected(ctx):
    # more embedding dimensions than the batched segment solver carries: the
    # per-component loop runs, as the reference's does
    z = load_fixture("disconnected_200")
    ip, ix, w = z["w_indptr"], z["w_indices"], z["w_data"]
    out = si.spectral_init((ip, ix, w), 8, 42, data=z["raw_X"],
                           config=si.SpectralInitConfig(tol=1e-6), ctx=ctx)
    st, ref = O.spectral_init(ip, ix, w, 8, 42, raw_data=z["raw_X"], tol=1e-6)
    assert st == 0 and out.shape == ref.shape == (200, 8)
    assert abs(np.abs(out).max() - 10.0) < 1e-3
    np.testing.assert_allclose(out, ref, atol=2e-3)
