"""The reference's solver and output contracts, on the device path.

* adversarial graphs (tests/integration/test_adversarial_graphs.rs:395-575):
  barbell, path, star, epsilon-bridge, complete bipartite, ring, exponentially
  weighted path, single edge, K_20 and lollipop through si_spectral_init --
  shape, finite values, community separation along the Fiedler direction,
  coordinate stability, the TooFewNodes error; where the wanted eigenvalues
  are separated, the coordinates against the oracle pipeline (same noise
  stream);
* solver levels (:580-640): level 0 below the dense threshold, rSVD on the
  path P_2001;
* warm restarts of the plain LOBPCG on the ring C_2000 with seed 42
  (tests/integration/test_solver_invariants.rs:209-230);
* the property sweep over rings n = 50..200
  (tests/integration/test_rq4_property_tests.rs): sorted eigenvalues in
  [0, 2], lambda_0 ~ 0, residual < 1e-5, V^T V = I, finite output with
  non-degenerate columns, bitwise-identical reruns; plus the closed form
  1 - cos(2 pi j / n).
"""
import numpy as np
import pytest

import oracle as O
from helpers import path_graph, ring_eigenvalues, ring_graph

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


def _csr(rows, n):
    indptr = np.zeros(n + 1, np.int64)
    indices, data = [], []
    for i in range(n):
        for j, w in sorted(rows[i]):
            indices.append(j)
            data.append(w)
        indptr[i + 1] = len(indices)
    return indptr, np.array(indices, np.int32), np.array(data, np.float32)


def complete(n):
    return _csr([[(j, 1.0) for j in range(n) if j != i] for i in range(n)], n)


def epsilon_bridge(cs, bw):
    rows = [[(j, 1.0) for j in range(cs) if j != i] for i in range(cs)]
    rows += [[(j, 1.0) for j in range(cs, 2 * cs) if j != i] for i in range(cs, 2 * cs)]
    rows[cs - 1].append((cs, bw))
    rows[cs].append((cs - 1, bw))
    return _csr(rows, 2 * cs)


def star(n):
    rows = [[(j, 1.0) for j in range(1, n)]] + [[(0, 1.0)] for _ in range(1, n)]
    return _csr(rows, n)


def complete_bipartite(m, k):
    rows = [[(j, 1.0) for j in range(m, m + k)] for _ in range(m)]
    rows += [[(j, 1.0) for j in range(m)] for _ in range(k)]
    return _csr(rows, m + k)


def exponential_path(n):
    w = [np.float32(10.0) ** np.float32(-8.0 * e / (n - 1)) for e in range(n - 1)]
    rows = [[] for _ in range(n)]
    for e in range(n - 1):
        rows[e].append((e + 1, float(w[e])))
        rows[e + 1].append((e, float(w[e])))
    return _csr(rows, n)


def lollipop(cs, pl):
    n = cs + pl
    rows = [[(j, 1.0) for j in range(cs) if j != i] for i in range(cs)]
    rows += [[] for _ in range(pl)]
    rows[cs - 1].append((cs, 1.0))
    for i in range(cs, n):
        rows[i].append((i - 1, 1.0))
        if i + 1 < n:
            rows[i].append((i + 1, 1.0))
    return _csr(rows, n)


GRAPHS = {
    "barbell20": lambda: epsilon_bridge(20, 1.0),
    "path100": lambda: path_graph(100),
    "star50": lambda: star(50),
    "epsilon_bridge20": lambda: epsilon_bridge(20, 1e-6),
    "complete_bipartite10_10": lambda: complete_bipartite(10, 10),
    "ring100": lambda: ring_graph(100),
    "exponential_path100": lambda: exponential_path(100),
    "complete20": lambda: complete(20),
    "lollipop10_20": lambda: lollipop(10, 20),
}


def _separated(g, k=3):
    """True when the k smallest eigenvalues of the normalized Laplacian are
    simple (so the embedding is unique up to the fixed sign rule)."""
    L, _, _ = O.laplacian_from_graph(*g)
    lam = np.linalg.eigvalsh(O.spmm(L, np.eye(g[0].size - 1)))
    gaps = np.diff(lam[:k + 1])
    return bool(np.all(gaps > 1e-6 * np.maximum(lam[1:k + 1], 1e-3)))


def _assert_same_up_to_tied_signs(out, ref, atol):
    """Column by column equal, or equal after a flip where the sign rule has a
    tie: normalize_signs makes the largest-|x| entry positive (scaling.rs:62-74),
    and a mirror-symmetric graph's antisymmetric eigenvector has two largest
    entries of equal size and opposite sign, so its sign is decided by rounding."""
    for c in range(out.shape[1]):
        if np.allclose(out[:, c], ref[:, c], atol=atol):
            continue
        mag = np.sort(np.abs(ref[:, c]))[::-1]
        tied = mag[0] - mag[1] < 1e-3 * mag[0]
        assert tied and np.allclose(out[:, c], -ref[:, c], atol=atol), c


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_adversarial_graph_valid_embedding(ctx, name):
    print(f"case adversarial {name}", flush=True)
    g = GRAPHS[name]()
    n = g[0].size - 1
    out = si.spectral_init(g, 2, 42, ctx=ctx)
    assert out.shape == (n, 2) and out.dtype == np.float32
    assert np.isfinite(out).all()
    assert abs(np.abs(out).max() - 10.0) < 1e-3
    if _separated(g):
        st, ref = O.spectral_init(*g, 2, 42)
        assert st == 0
        _assert_same_up_to_tied_signs(out, ref, 2e-3)


@pytest.mark.parametrize("name", ["barbell20", "epsilon_bridge20"])
def test_cliques_separate_along_the_fiedler_direction(ctx, name):
    out = si.spectral_init(GRAPHS[name](), 2, 42, ctx=ctx)
    c0, c1 = out[:20, 0], out[20:, 0]
    gap = abs(c0.mean() - c1.mean())
    intra = max(np.ptp(c0), np.ptp(c1))
    assert gap > intra, (gap, intra)


@pytest.mark.parametrize("name", ["star50", "complete_bipartite10_10"])
def test_coordinate_stability(ctx, name):
    out = si.spectral_init(GRAPHS[name](), 2, 42, ctx=ctx)
    assert np.abs(out).max() < 20.0


def test_single_edge_too_few_nodes(ctx):
    g = _csr([[(1, 1.0)], [(0, 1.0)]], 2)
    with pytest.raises(si.TooFewNodes):
        si.spectral_init(g, 2, 42, ctx=ctx)


def test_level0_for_small_graphs_and_rsvd_on_long_path(ctx):
    lap = si.Laplacian(ctx, *path_graph(6))
    ev, vec, rep = si.solve_eigenproblem(lap, 2, 42)
    assert rep.level == 0
    np.testing.assert_allclose(ev, 1.0 - np.cos(np.pi * np.arange(3) / 5), atol=1e-12)
    # test_level_3_rsvd_valid_on_large_path: rSVD straight on P_2001
    lap = si.Laplacian(ctx, *path_graph(2001))
    ev, vec, rep = si.rsvd_solve(lap, 2, 42)
    assert ev.shape == (3,) and np.all(ev >= -1e-3) and abs(ev[0]) < 0.1
    assert torch.isfinite(vec).all()


def test_ring2000_plain_lobpcg_warm_restart_fires(ctx):
    # test_solver_invariants.rs:209-230: ring C_2000, seed 42 must fire a warm
    # restart.  The ring's nontrivial eigenvalues come in exactly degenerate
    # pairs 1.5e-5 apart from the next pair, and the first pass (300
    # iterations, lobpcg.rs:315-320) does not reach the gate: the solve
    # continues from the partial eigenvectors (lobpcg.rs:373-394).  The
    # reference's restart_count counts only the other kind, a pass that met
    # linfa's rnorm gate but failed the residual after Rayleigh-Ritz
    # (lobpcg.rs:429-450); whether linfa's rnorm lets that happen is specific
    # to its iteration and is not reproduced here, so `passes` records every
    # warm restart and `restarts` that kind alone.
    print("case ring2000 restarts", flush=True)
    ip, ix, w = ring_graph(2000)
    lap = si.Laplacian(ctx, ip, ix, w)
    ev, vec, rep = si.lobpcg_solve(lap, 2, 42)
    print(f"  passes {rep.passes} restarts {rep.restarts} iterations {rep.iterations} "
          f"residual {rep.max_residual:.2e}", flush=True)
    assert rep.max_residual < 1e-5
    np.testing.assert_allclose(ev, ring_eigenvalues(2000, 3), atol=1e-7)
    assert rep.passes >= 2
    assert rep.iterations > 300


def test_per_level_timing(ctx):
    # [timing:level_N] (solvers/mod.rs:76-223): the levels tried carry their time
    z_ip, z_ix, z_w = path_graph(6)
    lap = si.Laplacian(ctx, z_ip, z_ix, z_w)
    ev, vec, rep = si.solve_eigenproblem(lap, 2, 42)
    assert rep.level == 0 and rep.level_ms[0] > 0 and not any(rep.level_ms[1:])
    ip, ix, w = ring_graph(3000)
    lap = si.Laplacian(ctx, ip, ix, w)
    ev, vec, rep = si.solve_eigenproblem(lap, 2, 42, si.SpectralInitConfig(tol=1e-6))
    assert rep.level >= 1
    assert rep.level_ms[0] == 0 and all(rep.level_ms[lv] > 0 for lv in range(1, rep.level + 1))
    assert not any(rep.level_ms[rep.level + 1:])


def test_ring_property_sweep(ctx):
    # test_rq4_property_tests.rs over every n in [50, 200]
    for n in range(50, 201):
        ip, ix, w = ring_graph(n)
        lap = si.Laplacian(ctx, ip, ix, w)
        ev, vec, rep = si.solve_eigenproblem(lap, 2, 42)
        v = vec.cpu().numpy()
        assert np.all(np.diff(ev) >= -1e-12), n
        assert np.all(ev >= -1e-12) and np.all(ev <= 2.0 + 1e-12), n
        assert abs(ev[0]) < 1e-6, n
        np.testing.assert_allclose(ev, ring_eigenvalues(n, 3), atol=1e-10)
        L, _, _ = O.laplacian_from_graph(ip, ix, w)
        assert O.max_eigenpair_residual(L, ev, v) < 1e-5, n
        assert np.linalg.norm(v.T @ v - np.eye(3)) < 1e-8, n
        out1 = si.spectral_init((ip, ix, w), 2, 42, ctx=ctx)
        out2 = si.spectral_init((ip, ix, w), 2, 42, ctx=ctx)
        assert np.isfinite(out1).all(), n
        assert np.all(out1.var(axis=0) > 1e-10), n
        np.testing.assert_array_equal(out1, out2)
