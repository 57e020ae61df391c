"""GPU tests of the shift-and-invert LOBPCG (escalation level 2,
src/solvers/sinv.rs:106-189) through si_lobpcg_sinv_solve.

The reference's own bar (tests/integration/test_comp_g_sinv.rs:36-93): on every
connected fixture, eigenvalues within 1e-8 of the fixture eigenvalues and
max residual ||Lv - lambda v|| below 1e-8; disconnected fixtures are skipped
there and here.  Ring C_2500 must give non-zero eigenvectors
(test_solver_invariants.rs:255-270); rings and paths are also checked against
their closed-form spectra, including one stored in reverse Cuthill-McKee order
(n >= 4096, bandwidth 2).  The chain test forces level 1 to fail (one filter
round) and checks that level 2 produces the result.  Graphs with no narrow band
(a random graph) take the dense Cholesky fallback up to 65536 nodes, checked
against scipy's direct solve and the oracle's LOBPCG.
"""
import numpy as np
import pytest

import oracle as O
from conftest import load_fixture
from helpers import (chained_random_graphs, path_graph, random_symmetric_graph, ring_eigenvalues,
                     ring_graph)

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


def _residual(ip, ix, w, ev, vec):
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    return O.max_eigenpair_residual(L, np.asarray(ev), vec.cpu().numpy())


@pytest.mark.parametrize("name", ["near_dupes_100", "blobs_connected_200", "circles_300",
                                  "moons_200", "blobs_connected_2000"])
def test_sinv_matches_fixture_eigenvalues(ctx, name):
    z = load_fixture(name)
    assert int(z["comp_n"]) == 1
    ip, ix, w = z["w_indptr"], z["w_indices"], z["w_data"]
    lap = si.Laplacian(ctx, ip, ix, w)
    print(f"sinv fixture {name} n={lap.n}", flush=True)
    ev, vec, rep = si.lobpcg_sinv_solve(lap, 2, 42)
    ref = np.asarray(z["eig_values"])
    k = min(len(ev), len(ref))
    np.testing.assert_allclose(ev[:k], ref[:k], rtol=0, atol=1e-8)
    assert np.all(np.diff(ev) >= -1e-12)
    res = _residual(ip, ix, w, ev, vec)
    assert res < 1e-8, res
    assert rep.level == 2 and rep.max_residual < 1e-8


def test_sinv_ring_2500_non_zero_and_closed_form(ctx):
    n = 2500
    ip, ix, w = ring_graph(n)
    lap = si.Laplacian(ctx, ip, ix, w)
    ev, vec, rep = si.lobpcg_sinv_solve(lap, 2, 42)
    v = vec.cpu().numpy()
    assert np.all(np.linalg.norm(v, axis=0) > 1e-6)
    np.testing.assert_allclose(ev, ring_eigenvalues(n, 3), rtol=0, atol=1e-8)
    assert _residual(ip, ix, w, ev, vec) < 1e-8


@pytest.mark.parametrize("kind,n", [("ring", 6000), ("path", 5000)])
def test_sinv_locality_ordered_band(ctx, kind, n):
    # n >= 4096: stored in reverse Cuthill-McKee order, bandwidth 1 (path) or 2 (ring)
    ip, ix, w = ring_graph(n) if kind == "ring" else path_graph(n)
    lap = si.Laplacian(ctx, ip, ix, w)
    print(f"sinv {kind} n={n}", flush=True)
    ev, vec, rep = si.lobpcg_sinv_solve(lap, 2, 42)
    if kind == "ring":
        ref = ring_eigenvalues(n, 3)
    else:
        ref = 1.0 - np.cos(np.pi * np.arange(3) / (n - 1))
    np.testing.assert_allclose(ev, ref, rtol=0, atol=1e-9)
    assert _residual(ip, ix, w, ev, vec) < 1e-8


def test_sinv_returns_none_when_band_too_large(ctx):
    # a random graph keeps a wide band under any ordering: sinv.rs returns None
    # when its factorisation is unavailable, here when the band exceeds the cap
    # and n is above the dense fallback's limit (65536)
    ip, ix, w = random_symmetric_graph(70000, 8, seed=11)
    lap = si.Laplacian(ctx, ip, ix, w)
    with pytest.raises(si.ConvergenceFailure):
        si.lobpcg_sinv_solve(lap, 2, 42)


@pytest.mark.parametrize("kind", ["ring", "fixture"])
def test_dense_factor_solve_matches_scipy(ctx, kind):
    # the dense fallback (rocSOLVER Cholesky, two right-side TRSMs on the
    # row-major block), forced on graphs that also have a narrow band
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    if kind == "fixture":
        z = load_fixture("blobs_connected_2000")
        ip, ix, w = z["w_indptr"], z["w_indices"], z["w_data"]
    else:
        ip, ix, w = ring_graph(2500)
    lap = si.Laplacian(ctx, ip, ix, w)
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    nn = lap.n
    m = sp.csr_matrix((L[2], L[1], L[0]), shape=(nn, nn)) + 1e-4 * sp.eye(nn)
    x = torch.from_numpy(np.random.default_rng(nn + 1).standard_normal((nn, 3))).cuda()
    si.sinv_dense(True)
    try:
        y, bw = si.band_solve(lap, x, 1e-4)
    finally:
        si.sinv_dense(False)
    assert bw == nn - 1  # the dense factor is the full lower triangle
    ref = sla.spsolve(m.tocsc(), x.cpu().numpy())
    err = np.abs(y.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < 1e-10, err
    y_band, _ = si.band_solve(lap, x, 1e-4)  # the banded factor (the whole triangle below 4096 nodes)
    assert np.abs(y.cpu().numpy() - y_band.cpu().numpy()).max() / np.abs(ref).max() < 1e-10


def test_sinv_dense_fallback_on_wide_band_graph(ctx):
    # three random graphs of 12000 nodes in a weak chain: the band exceeds the
    # banded limits, so level 2 factorises densely (10.4 GB) instead of
    # returning None; eigenpairs against the oracle's LOBPCG and the
    # reference's 1e-8 residual bar
    ip, ix, w = chained_random_graphs(12000, 3, seed=5)
    lap = si.Laplacian(ctx, ip, ix, w)
    x = torch.from_numpy(np.random.default_rng(9).standard_normal((lap.n, 2))).cuda()
    _, bw = si.band_solve(lap, x, 1e-4)
    assert bw == lap.n - 1, bw  # the automatic choice took the dense factor
    print(f"sinv dense fallback n={lap.n}", flush=True)
    ev, vec, rep = si.lobpcg_sinv_solve(lap, 2, 42)
    assert rep.level == 2
    L, _, sq = O.laplacian_from_graph(ip, ix, w)
    ref = O.lobpcg_solve(L, 2, 42, sq, tol=1e-10)
    assert ref is not None and ref["converged"]
    np.testing.assert_allclose(ev, ref["evals"], rtol=1e-6, atol=1e-12)
    assert _residual(ip, ix, w, ev, vec) < 1e-8


def test_sinv_deterministic(ctx):
    z = load_fixture("blobs_connected_200")
    lap = si.Laplacian(ctx, z["w_indptr"], z["w_indices"], z["w_data"])
    ev1, v1, _ = si.lobpcg_sinv_solve(lap, 2, 42)
    ev2, v2, _ = si.lobpcg_sinv_solve(lap, 2, 42)
    assert np.array_equal(ev1, ev2)
    assert torch.equal(v1, v2)


def test_chain_escalates_to_level_2(ctx):
    # level 0 off (dense threshold 0) and level 1 capped at one filter round, so
    # the chain reaches the shift-and-invert solve (solvers/mod.rs:131-153)
    n = 5000
    ip, ix, w = ring_graph(n)
    lap = si.Laplacian(ctx, ip, ix, w)
    cfg = si.SpectralInitConfig(tol=1e-6, dense_threshold=0, max_outer=1)
    ev, vec, rep = si.solve_eigenproblem(lap, 2, 42, cfg)
    assert rep.level == 2, rep.level
    np.testing.assert_allclose(ev, ring_eigenvalues(n, 3), rtol=0, atol=1e-9)
    assert rep.max_residual < 1e-6


@pytest.mark.parametrize("kind,n", [("ring", 2500), ("ring", 6000), ("path", 5000), ("fixture", 0)])
def test_band_solve_matches_scipy(ctx, kind, n):
    # the operator of sinv.rs:139-153: (L + 1e-4 I)^-1 applied to a block,
    # against scipy's sparse direct solve on the oracle's FP64 Laplacian
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    if kind == "fixture":
        z = load_fixture("blobs_connected_2000")
        ip, ix, w = z["w_indptr"], z["w_indices"], z["w_data"]
    else:
        ip, ix, w = ring_graph(n) if kind == "ring" else path_graph(n)
    lap = si.Laplacian(ctx, ip, ix, w)
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    nn = lap.n
    m = sp.csr_matrix((L[2], L[1], L[0]), shape=(nn, nn)) + 1e-4 * sp.eye(nn)
    x = torch.from_numpy(np.random.default_rng(nn).standard_normal((nn, 3))).cuda()
    y, bw = si.band_solve(lap, x, 1e-4)
    ref = sla.spsolve(m.tocsc(), x.cpu().numpy())
    # cond(M) <= 2e4: a backward-stable solve is accurate to ~cond * eps
    err = np.abs(y.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < 1e-10, (bw, err)
    if kind == "path":
        assert bw == 1
    if kind == "ring" and n >= 4096:
        assert bw == 2
