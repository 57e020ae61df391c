"""The five benchmark configurations at their stated parameters, on the GPU.

C1  n = 2000, 10 equal clusters in 50-d, exact 15-NN fuzzy graph, FP64,
    n_components = 2: device dense EVD (level 0), the FP32-bulk solver and the
    plain LOBPCG against numpy's dense eigh and the oracle's LOBPCG.  Also the
    configuration's literal N(0, 10^2 I) centres, which split the graph into
    one component per cluster: spectral_init's multi-component path against the
    oracle pipeline.
C2  n = 100 000, 20 Dirichlet(2) clusters in 50-d, block 8, Chebyshev degree
    8, tol 1e-6: both device solvers against the oracle's FP64 LOBPCG.
C3  n = 1 000 000, 40 Zipf(1.2) clusters + 5 % background in 30-d: at full
    size through properties -- an independent FP64 residual (device and CPU
    oracle), ascending eigenvalues in [0, 2], lambda_0 ~ 0 with v_0 along
    D^1/2 1, orthonormal vectors, agreement with the plain LOBPCG within the
    Davis-Kahan bound of their residuals.
C4  64 disconnected components, sizes log-uniform in [50, 200 000]:
    spectral_init with the batched device solve against the per-component loop.
C5  the learned warm start on a held-out graph of the training construction
    scaled to n = 1 000 000: the residual-judged choice, and warm-started
    solves that meet the gate.

Parity bar: eigenvalues 1e-6 relative (absolute 1e-9 for lambda_0), subspace
angle below 1e-4 rad for gap-separated eigenvectors, FP64 residual <= tol.
The center spread of the connected workloads (graphs.CENTER_SCALE) keeps their
graphs connected; C1's literal spread is tested separately.
"""
import math
import os

import numpy as np
import pytest

import oracle as O
from helpers import max_subspace_angle, principal_angles

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


def _workload(name, seed=0, n=None):
    import graphs
    w = graphs.workload(name, seed=seed, n=n, device="cuda")
    assert graphs.n_connected_components(w["indptr"], w["indices"]) == 1, name
    return w


def _check_pairs(ev, vec, ref_ev, ref_vec, L, tol_res, k):
    """eigenvalues, per-gap subspace angles and the FP64 residual."""
    full = np.asarray(ref_ev)
    ev = np.asarray(ev)[:k]
    assert abs(ev[0] - full[0]) < 1e-9
    np.testing.assert_allclose(ev[1:], full[1:k], rtol=1e-6)
    res = O.max_eigenpair_residual(L, ev, vec[:, :k])
    assert res <= tol_res, res
    for j in range(k):
        gaps = [abs(full[j] - full[i]) for i in range(len(full)) if i != j]
        if min(gaps) > 1e-3 * max(abs(full[j]), 1e-3):
            ang = max_subspace_angle(vec[:, [j]], ref_vec[:, [j]])
            assert ang < 1e-4, (j, ang)


# ------------------------------------------------------------------ C1
@pytest.fixture(scope="module")
def c1():
    w = _workload("blobs2k")
    L, _, sq = O.laplacian_from_graph(w["indptr"], w["indices"], w["data"])
    lam, vec = np.linalg.eigh(O.spmm(L, np.eye(w["n"])))
    return w, L, sq, lam[:8], vec[:, :8]


def test_c1_dense_evd_level0(ctx, c1):
    print("case C1 dense", flush=True)
    w, L, sq, lam, vec = c1
    lap = si.Laplacian(ctx, w["indptr"], w["indices"], w["data"])
    ev, v = si.dense_evd(lap, 3)
    np.testing.assert_allclose(ev, lam[:3], atol=1e-9)
    _check_pairs(ev, v.cpu().numpy(), lam, vec, L, 1e-10, 3)
    # the chain's level 0 below the dense threshold (solvers/mod.rs:83-109)
    ev0, v0, rep = si.solve_eigenproblem(lap, 2, 42, si.SpectralInitConfig(tol=1e-6,
                                                                          dense_threshold=2001))
    assert rep.level == 0
    _check_pairs(ev0, v0.cpu().numpy(), lam, vec, L, 1e-10, 3)


def test_c1_lobpcg_solvers(ctx, c1):
    print("case C1 lobpcg", flush=True)
    w, L, sq, lam, vec = c1
    lap = si.Laplacian(ctx, w["indptr"], w["indices"], w["data"])
    ev, v, rep = si.solve_eigenproblem(lap, 2, 42, si.SpectralInitConfig(tol=1e-6))
    assert rep.level == 1
    _check_pairs(ev, v.cpu().numpy(), lam, vec, L, 1e-6, 3)
    ev, v, rep = si.lobpcg_solve(lap, 2, 42, tol=1e-6)
    _check_pairs(ev, v.cpu().numpy(), lam, vec, L, 1e-6, 3)
    # the oracle's restatement of the reference path on the same graph
    r = O.lobpcg_solve(L, 2, 42, sq, tol=1e-6)
    _check_pairs(r["evals"], r["evecs"], lam, vec, L, 1e-6, 3)


def test_c1_literal_centres_take_the_multi_component_path(ctx):
    import graphs
    print("case C1 literal centres", flush=True)
    x, _ = graphs.gmm_points(2000, 50, np.ones(10), 0, center_scale=10.0)
    ip, ix, w = graphs.fuzzy_knn_graph(x, 15, device="cuda")
    labels, nc = O.find_components(ip, ix)
    assert nc == 10
    out = si.spectral_init((ip, ix, w), 2, 42, data=x, config=si.SpectralInitConfig(tol=1e-6),
                           ctx=ctx)
    st, ref = O.spectral_init(ip, ix, w, 2, 42, raw_data=x, tol=1e-6)
    assert st == 0 and out.shape == ref.shape == (2000, 2)
    assert abs(np.abs(out).max() - 10.0) < 1e-3
    np.testing.assert_allclose(out, ref, atol=2e-3)


# ------------------------------------------------------------------ C2
@pytest.fixture(scope="module")
def c2():
    w = _workload("gmm100k")
    L, _, sq = O.laplacian_from_graph(w["indptr"], w["indices"], w["data"])
    import time
    t = time.perf_counter()
    O.set_threads(16)
    # 24 columns (more than the 20 clusters), low-pass filtered twice with
    # degree 24 on [0.1, 2] and orthonormalised after each: the wanted pairs
    # then converge at the rate of the gap to the bulk
    x0 = O.normal_block(42, 1, w["n"] * 24).reshape(w["n"], 24)
    x0[:, 0] = sq / np.linalg.norm(sq)
    for _ in range(2):
        x0, _ = np.linalg.qr(O.chebyshev_filter(L, x0, 0.1, 2.0, 24))
    ref = O.lobpcg_core(L, x0, tol=1e-9, maxiter=3000)
    O.set_threads(1)
    print(f"  oracle FP64 LOBPCG (24 columns): {ref['iterations']} iterations, "
          f"{time.perf_counter() - t:.1f} s", flush=True)
    assert ref["converged"]
    return w, L, ref


def test_c2_block8_degree8_both_solvers(ctx, c2):
    print("case C2 100K block 8", flush=True)
    w, L, ref = c2
    lap = si.Laplacian(ctx, w["indptr"], w["indices"], w["data"])
    cfg = si.SpectralInitConfig(tol=1e-6, block_size=8, cheb_degree=8)
    ev, v, rep = si.solve_eigenproblem(lap, 2, 42, cfg)
    assert rep.level == 1 and rep.block_size == 8
    _check_pairs(ev, v.cpu().numpy(), ref["evals"], ref["evecs"], L, 1e-6, 3)
    ev, v, rep = si.lobpcg_solve(lap, 2, 42, tol=1e-6)
    _check_pairs(ev, v.cpu().numpy(), ref["evals"], ref["evecs"], L, 1e-6, 3)


# ------------------------------------------------------------------ C3
@pytest.fixture(scope="module")
def c3(ctx):
    w = _workload("gmm1m")
    lap = si.Laplacian(ctx, w["indptr"], w["indices"], w["data"])
    return w, lap


def test_c3_million_node_properties(ctx, c3):
    print("case C3 1M", flush=True)
    w, lap = c3
    ev, v, rep = si.solve_eigenproblem(lap, 2, 42, si.SpectralInitConfig(tol=1e-6))
    assert rep.level == 1 and rep.max_residual <= 1e-6
    # an independent FP64 residual: the device entry point and the CPU oracle
    # on the exported Laplacian, both from the returned pairs
    dres = si.max_eigenpair_residual(lap, ev, v)
    deg, sq, L = lap.export()
    vh = v.cpu().numpy()
    ores = O.max_eigenpair_residual(L, ev, vh)
    assert dres <= 1e-6 and ores <= 1e-6
    assert dres == pytest.approx(ores, rel=1e-9)
    assert rep.max_residual == pytest.approx(ores, rel=1e-6)
    # ascending in [0, 2], the trivial pair first
    assert np.all(np.diff(ev) >= 0) and ev[0] >= -1e-12 and ev[-1] <= 2.0
    assert abs(ev[0]) < 1e-9
    assert max_subspace_angle(vh[:, [0]], sq[:, None]) < 1e-4
    np.testing.assert_allclose(vh.T @ vh, np.eye(3), atol=1e-8)
    # the plain LOBPCG: eigenvalues to 1e-6 and the pair's subspace within the
    # Davis-Kahan bound (||R_fast|| + ||R_plain||) / (lambda_4 - lambda_3)
    pev, pv, prep = si.lobpcg_solve(lap, 2, 42, tol=1e-6)
    assert prep.max_residual <= 1e-6
    assert abs(pev[0] - ev[0]) < 1e-9
    np.testing.assert_allclose(pev[1:], ev[1:], rtol=1e-6)
    ev4, _, rep4 = si.solve_eigenproblem(lap, 3, 42, si.SpectralInitConfig(tol=1e-6))
    gap = ev4[3] - ev4[2]
    bound = 2.0 * (rep.max_residual + prep.max_residual) / gap
    ang = principal_angles(vh, pv.cpu().numpy())[0]
    print(f"  lambda {ev} lambda_4 {ev4[3]:.6g} angle {ang:.2e} bound {bound:.2e}", flush=True)
    assert ang <= max(bound, 1e-6)


# ------------------------------------------------------------------ C4
def test_c4_sixty_four_components_batched_vs_loop(ctx):
    import graphs
    print("case C4 comp64", flush=True)
    w = graphs.workload("comp64", seed=0, device="cuda")
    ip, ix, wt, x, sizes = w["indptr"], w["indices"], w["data"], w["x"], w["sizes"]
    assert len(sizes) == 64 and min(sizes) >= 50 and max(sizes) <= 200_000
    labels, nc = O.find_components(ip, ix)
    assert nc == 64
    outs = {}
    for batch in (True, False):
        cfg = si.SpectralInitConfig(tol=1e-6, batch_components=batch)
        out, rep = si.spectral_init((ip, ix, wt), 2, 42, data=x, config=cfg, ctx=ctx,
                                    return_report=True)
        assert out.shape == (ip.size - 1, 2) and np.isfinite(out).all()
        assert abs(np.abs(out).max() - 10.0) < 1e-3
        assert rep.max_residual <= 1e-6
        outs[batch] = out
    print(f"  max |batched - loop| = {np.abs(outs[True] - outs[False]).max():.2e}", flush=True)
    np.testing.assert_allclose(outs[True], outs[False], atol=1e-3)


# ------------------------------------------------------------------ C5
def test_c5_warm_start_held_out_million_nodes(ctx):
    import graphs
    print("case C5 warm start 1M", flush=True)
    path = os.path.join(os.path.dirname(si.__file__), "warmstart_mpnn.npy")
    net = si.WarmStartNet(ctx, params=np.load(path))
    wl = graphs.workload("gmm100k", seed=6000, n=1_000_000, device="cuda")
    ip, ix, wt, _ = graphs.largest_connected(wl["indptr"], wl["indices"], wl["data"])
    lap = si.Laplacian(ctx, ip, ix, wt)
    blk, used, res_net, res_rand = net.warm_start(lap, seed=42)
    assert used == (res_net < res_rand)  # random start whenever the net's residual is worse
    assert math.isfinite(res_net) and math.isfinite(res_rand)
    plain = si.SpectralInitConfig(tol=1e-6, solver=si.Solver.PlainLobpcg)
    _, _, r_rand = si.solve_eigenproblem(lap, 2, 42, plain)
    plain_w = si.SpectralInitConfig(tol=1e-6, solver=si.Solver.PlainLobpcg,
                                    warm_start=blk.data_ptr())
    evw, vw, r_warm = si.solve_eigenproblem(lap, 2, 42, plain_w)
    fast_w = si.SpectralInitConfig(tol=1e-6, warm_start=blk.data_ptr())
    evf, vf, r_fast = si.solve_eigenproblem(lap, 2, 42, fast_w)
    print(f"  plain iterations random {r_rand.iterations} warm {r_warm.iterations}; "
          f"net used {used} ({res_net:.3e} vs {res_rand:.3e}); fast warm level {r_fast.level}",
          flush=True)
    for r in (r_rand, r_warm, r_fast):
        assert r.level == 1 and r.max_residual <= 1e-6
    assert si.max_eigenpair_residual(lap, evw, vw) <= 1e-6
    assert si.max_eigenpair_residual(lap, evf, vf) <= 1e-6
    np.testing.assert_allclose(evw[1:], evf[1:], rtol=1e-6)
