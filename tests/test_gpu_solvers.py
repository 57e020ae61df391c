"""GPU parity of the eigensolvers against the FP64 oracle (through the C ABI).

Parity bar (the eigensolver contract): eigenvalues to 1e-6 relative (absolute
for the trivial lambda_0 ~ 0), eigenvector subspace angle below 1e-4 rad for
gap-separated eigenvalues, FP64 residual ||Lv - lambda v|| / ||v|| below the
requested tolerance.
"""
import numpy as np
import pytest

import oracle as O
from conftest import load_fixture
from helpers import max_subspace_angle, principal_angles, ring_eigenvalues, ring_graph

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


def _graph(name):
    z = load_fixture(name)
    return z["w_indptr"], z["w_indices"], z["w_data"], z


@pytest.fixture(scope="module")
def gmm20k():
    import graphs
    w = graphs.workload("gmm100k", seed=3, n=20000)
    L, deg, sq = O.laplacian_from_graph(w["indptr"], w["indices"], w["data"])
    x0 = O.normal_block(42, 1, w["n"] * 16).reshape(w["n"], 16)
    x0[:, 0] = sq / np.linalg.norm(sq)
    ref = O.lobpcg_core(L, x0, tol=1e-9, maxiter=2000)
    assert ref["converged"]
    return w, L, ref


def _check_pairs(ev, vec, ref_ev, ref_vec, L, tol_res, k):
    """eigenvalues, per-gap subspace angles and FP64 residual."""
    full = np.asarray(ref_ev)
    ev = np.asarray(ev)[:k]
    ref_ev = full[:k]
    assert abs(ev[0] - ref_ev[0]) < 1e-9
    np.testing.assert_allclose(ev[1:], ref_ev[1:], rtol=1e-6)
    res = O.max_eigenpair_residual(L, ev, vec[:, :k])
    assert res <= tol_res, res
    # compare each eigenvector whose eigenvalue is separated from its neighbours
    for j in range(k):
        gaps = [abs(full[j] - full[i]) for i in range(len(full)) if i != j]
        if min(gaps) > 1e-3 * max(abs(full[j]), 1e-3):
            ang = max_subspace_angle(vec[:, [j]], ref_vec[:, [j]])
            assert ang < 1e-4, (j, ang)


# ------------------------------------------------------------------ plain LOBPCG (baseline)
@pytest.mark.parametrize("name", ["blobs_connected_2000", "blobs_5000", "blobs_connected_200",
                                  "circles_300"])
def test_plain_lobpcg_matches_scipy_fixture(ctx, name):
    # test_comp_f_lobpcg.rs:22-48 on the device path: eigenvalues vs scipy 1e-6, residual 1e-5
    print(f"case plain lobpcg {name}", flush=True)
    ip, ix, w, z = _graph(name)
    lap = si.Laplacian(ctx, ip, ix, w)
    ev, vec, rep = si.lobpcg_solve(lap, 2, 42)
    np.testing.assert_allclose(np.maximum(ev, 0), z["eig_values"], atol=1e-6)
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    assert O.max_eigenpair_residual(L, ev, vec.cpu().numpy()) < 1e-5
    assert rep.max_residual < 1e-5


def test_plain_lobpcg_tracks_the_oracle_iterate(ctx):
    # same seed -> same Philox start block and ChFSI filter -> the same
    # converged eigenpairs as the oracle's lobpcg_solve
    ip, ix, w, z = _graph("blobs_connected_2000")
    lap = si.Laplacian(ctx, ip, ix, w)
    L, _, sq = O.laplacian_from_graph(ip, ix, w)
    ev, vec, rep = si.lobpcg_solve(lap, 2, 42, tol=1e-8)
    ref = O.lobpcg_solve(L, 2, 42, sq, tol=1e-8)
    _check_pairs(ev, vec.cpu().numpy(), ref["evals"], ref["evecs"], L, 1e-8, 3)
    assert abs(rep.iterations - ref["iterations"]) <= max(5, ref["iterations"] // 5)


def test_plain_lobpcg_ring_closed_form(ctx):
    ip, ix, w = ring_graph(1200)
    lap = si.Laplacian(ctx, ip, ix, w)
    ev, vec, rep = si.lobpcg_solve(lap, 2, 42, tol=1e-9)
    np.testing.assert_allclose(ev, ring_eigenvalues(1200, 3), atol=5e-12)


def test_plain_lobpcg_regularized_level(ctx):
    ip, ix, w, z = _graph("blobs_connected_2000")
    lap = si.Laplacian(ctx, ip, ix, w)
    ev, vec, rep = si.lobpcg_solve(lap, 2, 42, regularize=True)
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    assert O.max_eigenpair_residual(L, ev, vec.cpu().numpy()) < 1e-5


# ------------------------------------------------------------------ FP32-bulk solver
@pytest.mark.parametrize("name", ["blobs_connected_2000", "blobs_connected_200", "moons_200",
                                  "circles_300"])
def test_fast_solver_matches_fixture_and_oracle(ctx, name):
    print(f"case fast {name}", flush=True)
    ip, ix, w, z = _graph(name)
    lap = si.Laplacian(ctx, ip, ix, w)
    ev, vec, rep = si.fast_solve(lap, 2, 42, si.SpectralInitConfig(tol=1e-6))
    L, _, sq = O.laplacian_from_graph(ip, ix, w)
    n = ip.size - 1
    dense = np.linalg.eigh(O.spmm(L, np.eye(n)) if n <= 2000 else None)
    _check_pairs(ev, vec.cpu().numpy(), dense[0][:8], dense[1][:, :8], L, 1e-6, 3)
    np.testing.assert_allclose(np.maximum(ev, 0), z["eig_values"], atol=1e-6)


def test_fast_solver_gmm20k_matches_oracle(ctx, gmm20k):
    w, L, ref = gmm20k
    print("case fast gmm20k", flush=True)
    lap = si.Laplacian(ctx, w["indptr"], w["indices"], w["data"])
    ev, vec, rep = si.fast_solve(lap, 2, 42, si.SpectralInitConfig(tol=1e-6))
    v = vec.cpu().numpy()
    _check_pairs(ev, v, ref["evals"], ref["evecs"], L, 1e-6, 3)
    assert rep.max_residual <= 1e-6


@pytest.mark.parametrize("block", [16, 32, 64])
def test_fast_solver_block_widths(ctx, gmm20k, block):
    w, L, ref = gmm20k
    lap = si.Laplacian(ctx, w["indptr"], w["indices"], w["data"])
    ev, vec, rep = si.fast_solve(lap, 2, 7, si.SpectralInitConfig(tol=1e-6, block_size=block))
    assert rep.block_size == block
    _check_pairs(ev, vec.cpu().numpy(), ref["evals"], ref["evecs"], L, 1e-6, 3)


def test_fast_solver_deterministic(ctx):
    ip, ix, w, z = _graph("blobs_connected_2000")
    lap = si.Laplacian(ctx, ip, ix, w)
    ev1, v1, _ = si.fast_solve(lap, 2, 42)
    ev2, v2, _ = si.fast_solve(lap, 2, 42)
    np.testing.assert_array_equal(ev1, ev2)
    assert torch.equal(v1, v2)


# ------------------------------------------------------------------ rSVD / dense / chain
def test_rsvd_ring6_closed_form(ctx):
    ip, ix, w = ring_graph(6)
    lap = si.Laplacian(ctx, ip, ix, w)
    ev, vec, rep = si.rsvd_solve(lap, 2, 42)
    np.testing.assert_allclose(ev, [0.0, 0.5, 0.5], atol=1e-10)
    assert rep.max_residual < 1e-10


def test_rsvd_matches_oracle_same_sketch(ctx):
    ip, ix, w, z = _graph("blobs_connected_200")
    lap = si.Laplacian(ctx, ip, ix, w)
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    ev, vec, rep = si.rsvd_solve(lap, 2, 42, oversampling=30, nbiter=8)
    oev, ovec = O.rsvd_solve(L, 2, 42, oversampling=30, nbiter=8)
    np.testing.assert_allclose(ev, oev, atol=1e-9)
    ang = principal_angles(vec.cpu().numpy(), ovec)
    assert ang[0] < 1e-6


@pytest.mark.parametrize("name", ["blobs_connected_200", "blobs_connected_2000"])
def test_rsvd_oversampling10_4iter_matches_oracle(ctx, name):
    # the benchmark's comparison settings (oversampling 10 -> a 13-column sketch
    # stored 16 wide, 4 power iterations) against the oracle with the same sketch
    print(f"case rsvd p=10 q=4 {name}", flush=True)
    ip, ix, w, z = _graph(name)
    lap = si.Laplacian(ctx, ip, ix, w)
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    ev, vec, rep = si.rsvd_solve(lap, 2, 42, oversampling=10, nbiter=4)
    oev, ovec = O.rsvd_solve(L, 2, 42, oversampling=10, nbiter=4)
    assert rep.block_size == 13
    np.testing.assert_allclose(ev, oev, atol=1e-9)
    assert principal_angles(vec.cpu().numpy(), ovec)[0] < 1e-6
    assert rep.max_residual == pytest.approx(O.max_eigenpair_residual(L, oev, ovec), rel=1e-6)


def test_rsvd_reference_sketch_beyond_64_columns(ctx):
    # rsvd_k_sub (rsvd.rs:84-87): rank + max(n/10, rank, 5) = 203 columns at
    # n = 2000, with the reference's 2 power iterations (the chain's level 4)
    print("case rsvd k_sub 203", flush=True)
    ip, ix, w, z = _graph("blobs_connected_2000")
    lap = si.Laplacian(ctx, ip, ix, w)
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    ev, vec, rep = si.rsvd_solve(lap, 2, 42)
    assert rep.block_size == 203
    oev, ovec = O.rsvd_solve(L, 2, 42)
    np.testing.assert_allclose(ev, oev, atol=1e-9)
    assert principal_angles(vec.cpu().numpy(), ovec)[0] < 1e-6


def test_rsvd_eigenvector_distinctness(ctx):
    # test_solver_invariants.rs:234-254: the returned eigenvectors are not collinear
    ip, ix, w, z = _graph("blobs_5000")
    lap = si.Laplacian(ctx, ip, ix, w)
    ev, vec, rep = si.rsvd_solve(lap, 2, 42)
    assert rep.block_size == 503
    v = vec.cpu().numpy()
    cos = abs(v[:, 0] @ v[:, 1]) / max(np.linalg.norm(v[:, 0]) * np.linalg.norm(v[:, 1]), 1e-300)
    assert cos < 0.99
    assert np.all(np.diff(ev) >= -1e-12) and np.isfinite(ev).all()


def test_rsvd_one_spmm_launch_per_application(ctx):
    # a 13-column sketch is applied as one 16-wide launch (not 8 + 4 + 1)
    ip, ix, w, z = _graph("blobs_5000")
    lap = si.Laplacian(ctx, ip, ix, w)
    with si.SpmmProfiler(ctx) as p:
        ev, vec, rep = si.rsvd_solve(lap, 2, 42, oversampling=10, nbiter=4)
    # 4 iterations x 2 applications + Q + M Q inside the solve, then the
    # FP64 residual's own application
    assert rep.spmm_calls == 10
    assert p.result["launches"] == rep.spmm_calls + 1


def test_dense_evd_small_matches_numpy(ctx):
    ip, ix, w, z = _graph("blobs_50")
    lap = si.Laplacian(ctx, ip, ix, w)
    ev, vec = si.dense_evd(lap, 3)
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    oev, ovec = O.dense_evd(L, 3)
    np.testing.assert_allclose(ev, oev, atol=1e-12)
    assert max_subspace_angle(vec.cpu().numpy(), ovec) < 1e-8


@pytest.mark.parametrize("name", ["blobs_50", "blobs_connected_200", "blobs_connected_2000"])
def test_solve_eigenproblem_chain(ctx, name):
    ip, ix, w, z = _graph(name)
    lap = si.Laplacian(ctx, ip, ix, w)
    ev, vec, rep = si.solve_eigenproblem(lap, 2, 42, si.SpectralInitConfig(tol=1e-6))
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    q = O.max_eigenpair_residual(L, ev, vec.cpu().numpy())
    assert q < 1e-6 and rep.level in (0, 1)
    # solvers/mod.rs:32, 83-109: level 0 (dense EVD) below n = 2000
    assert rep.level == (0 if ip.size - 1 < 2000 else 1)
    np.testing.assert_allclose(np.maximum(ev, 0), z["eig_values"], atol=1e-6)


def test_solve_eigenproblem_plain_solver_option(ctx):
    ip, ix, w, z = _graph("blobs_connected_2000")
    lap = si.Laplacian(ctx, ip, ix, w)
    cfg = si.SpectralInitConfig(tol=1e-6, solver=si.Solver.PlainLobpcg)
    ev, vec, rep = si.solve_eigenproblem(lap, 2, 42, cfg)
    assert rep.level == 1 and rep.max_residual < 1e-6


# ------------------------------------------------------------------ spectral_init end to end
@pytest.mark.parametrize("name", ["blobs_connected_200", "blobs_connected_2000",
                                  "disconnected_200", "blobs_5000"])
def test_spectral_init_matches_oracle(ctx, name):
    print(f"case spectral_init {name}", flush=True)
    ip, ix, w, z = _graph(name)
    out, rep = si.spectral_init((ip, ix, w), 2, 42, data=z["raw_X"],
                                config=si.SpectralInitConfig(tol=1e-6), ctx=ctx,
                                return_report=True)
    st, ref = O.spectral_init(ip, ix, w, 2, 42, raw_data=z["raw_X"], tol=1e-6)
    assert st == 0
    assert out.shape == ref.shape and out.dtype == np.float32
    assert abs(np.abs(out).max() - 10.0) < 1e-3
    # same sign convention and the same noise stream: coordinates agree
    np.testing.assert_allclose(out, ref, atol=2e-3)


def test_spectral_init_errors(ctx):
    with pytest.raises(si.TooFewNodes):
        si.spectral_init((np.array([0, 1, 2]), np.array([1, 0]), np.ones(2)), 2, 0, ctx=ctx)
    with pytest.raises(si.InvalidGraph):
        si.spectral_init((np.array([0]), np.zeros(0), np.zeros(0)), 2, 0, ctx=ctx)
    z = load_fixture("disconnected_200")
    labels, nc = O.find_components(z["w_indptr"], z["w_indices"])
    if nc > 4:
        with pytest.raises(si.InvalidGraph):
            si.spectral_init((z["w_indptr"], z["w_indices"], z["w_data"]), 2, 0, ctx=ctx)
