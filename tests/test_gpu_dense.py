"""GPU parity of the device dense eigensolver (dense_evd, src/solvers/dense.rs:9-46)
for 64 < n <= 4096: Householder tridiagonalisation, Sturm multisection, inverse
iteration and back-transformation, called through si_dense_evd.

Bar (tests/test_comp_d_dense_evd.rs:7-66): eigenvalues within 1e-9 of the
FP64 oracle (numpy eigh on the same FP64 Laplacian), FP64 residual
||Lv - lambda v|| below 1e-10, orthonormal eigenvectors, and each
eigenvector within sin(angle) < 1e-12 + 1e-13 / gap of the oracle's (the
Davis-Kahan bound for residuals near 1e-13).  Degenerate
eigenvalues (the ring's pairs, the zero eigenvalue of a disconnected graph)
are checked through their invariant subspace.
"""
import numpy as np
import pytest

import oracle as O
from conftest import load_fixture
from helpers import path_graph, random_symmetric_graph, ring_eigenvalues, ring_graph

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


def _dense_l(ip, ix, w):
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    lip, lix, ld = L
    n = lip.size - 1
    a = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(lip))
    a[rows, lix] = ld
    return L, a


def _sin_angle(v, q):
    """sin of the largest principal angle between span(v) and span(q), both
    orthonormal: ||v - q q^T v||_2 (accurate down to eps, unlike arccos of
    the singular values, which bottoms out near 1.5e-8)."""
    return np.linalg.norm(v - q @ (q.T @ v), 2)


def _check(ev, vec, a, k, ev_tol=1e-9):
    vec = vec.cpu().numpy() if hasattr(vec, "cpu") else vec
    full, fvec = np.linalg.eigh(a)
    np.testing.assert_allclose(ev, full[:k], rtol=0, atol=ev_tol)
    assert np.all(np.diff(ev) >= 0)
    res = np.linalg.norm(a @ vec - vec * ev, axis=0).max()
    assert res < 1e-10, res
    orth = np.abs(vec.T @ vec - np.eye(k)).max()
    assert orth < 1e-10, orth
    # subspace of each eigenvalue cluster (gap 1e-6) against eigh's
    j = 0
    while j < k:
        e = j + 1
        while e < len(full) and full[e] - full[e - 1] < 1e-6:
            e += 1
        if e <= k:
            # Davis-Kahan: sin(angle) <= ||residual|| / gap, residuals ~1e-15
            gap = min(full[j] - full[j - 1] if j > 0 else np.inf,
                      full[e] - full[e - 1] if e < len(full) else np.inf)
            ang = _sin_angle(vec[:, j:e], fvec[:, j:e])
            assert ang < 1e-12 + 1e-13 / gap, (j, e, ang, gap)
        j = e


@pytest.mark.parametrize("name", ["near_dupes_100", "blobs_connected_200", "circles_300",
                                  "blobs_500", "blobs_connected_2000"])
def test_dense_evd_matches_eigh_on_fixtures(ctx, name):
    z = load_fixture(name)
    ip, ix, w = z["w_indptr"], z["w_indices"], z["w_data"]
    lap = si.Laplacian(ctx, ip, ix, w)
    _, a = _dense_l(ip, ix, w)
    print(f"dense_evd fixture {name} n={lap.n}", flush=True)
    ev, vec = si.dense_evd(lap, 4)
    _check(ev, vec, a, 4)


def test_dense_evd_disconnected_zero_multiplicity(ctx):
    z = load_fixture("disconnected_200")
    ip, ix, w = z["w_indptr"], z["w_indices"], z["w_data"]
    lap = si.Laplacian(ctx, ip, ix, w)
    _, a = _dense_l(ip, ix, w)
    ev, vec = si.dense_evd(lap, 5)
    _check(ev, vec, a, 5)
    # one (near-)zero eigenvalue per component; FP32 degrees leave them at
    # ~1e-8 of zero in the FP64 matrix itself (eigh agrees to 1e-9 above)
    ncomp = int(np.sum(np.abs(np.linalg.eigvalsh(a)) < 1e-6))
    assert ncomp >= 2 and np.all(np.abs(ev[:ncomp]) < 1e-6)


def test_dense_evd_ring_closed_form(ctx):
    n = 300
    ip, ix, w = ring_graph(n)
    lap = si.Laplacian(ctx, ip, ix, w)
    _, a = _dense_l(ip, ix, w)
    ev, vec = si.dense_evd(lap, 7)
    np.testing.assert_allclose(ev, ring_eigenvalues(n, 7), rtol=0, atol=1e-12)
    _check(ev, vec, a, 7)


def test_dense_evd_path_graph(ctx):
    n = 257
    ip, ix, w = path_graph(n)
    lap = si.Laplacian(ctx, ip, ix, w)
    _, a = _dense_l(ip, ix, w)
    ev, vec = si.dense_evd(lap, 6)
    # normalized Laplacian of P_n: 1 - cos(pi j / (n - 1))
    np.testing.assert_allclose(ev, 1.0 - np.cos(np.pi * np.arange(6) / (n - 1)), rtol=0, atol=1e-12)
    _check(ev, vec, a, 6)


@pytest.mark.parametrize("n,k", [(65, 3), (128, 3), (129, 5), (1024, 16), (4096, 8)])
def test_dense_evd_random_graphs(ctx, n, k):
    ip, ix, w = random_symmetric_graph(n, 8, seed=n)
    lap = si.Laplacian(ctx, ip, ix, w)
    _, a = _dense_l(ip, ix, w)
    print(f"dense_evd random n={n} k={k}", flush=True)
    ev, vec = si.dense_evd(lap, k)
    _check(ev, vec, a, k)


def test_dense_evd_deterministic(ctx):
    z = load_fixture("blobs_500")
    lap = si.Laplacian(ctx, z["w_indptr"], z["w_indices"], z["w_data"])
    ev1, v1 = si.dense_evd(lap, 3)
    ev2, v2 = si.dense_evd(lap, 3)
    assert np.array_equal(ev1, ev2)
    assert torch.equal(v1, v2)


def test_dense_evd_rejects_above_4096(ctx):
    ip, ix, w = ring_graph(4097)
    lap = si.Laplacian(ctx, ip, ix, w)
    with pytest.raises(si.SpectralError, match="4096"):
        si.dense_evd(lap, 3)


@pytest.mark.parametrize("name", ["blobs_connected_200", "blobs_500", "circles_300"])
def test_chain_takes_dense_level_below_threshold(ctx, name):
    # solvers/mod.rs:83-109: n < DENSE_N_THRESHOLD (2000) goes to level 0
    z = load_fixture(name)
    ip, ix, w = z["w_indptr"], z["w_indices"], z["w_data"]
    lap = si.Laplacian(ctx, ip, ix, w)
    ev, vec, rep = si.solve_eigenproblem(lap, 2, 42, si.SpectralInitConfig(tol=1e-6))
    assert rep.level == 0
    assert rep.max_residual < 1e-10
    L, a = _dense_l(ip, ix, w)
    _check(ev, vec, a, 3)
    if "eig_values" in z:
        np.testing.assert_allclose(np.maximum(ev, 0), z["eig_values"], atol=1e-6)
