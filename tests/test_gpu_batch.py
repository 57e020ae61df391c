"""Batched solve of the connected components of a disconnected graph.

si_spectral_init solves every component of at least 2*n_components nodes in
one device pass (one block-diagonal Laplacian; batched dense EVD for <= 64
nodes, the segmented FP32-bulk solver above) instead of the reference's
per-component loop (src/multi_component.rs:261-312).  Checked against the
per-component path, the FP64 oracle pipeline, and -- for the hipGraph-captured
rounds, planned on the device between rounds -- against the same solve run
eagerly, one round per host read.
"""
import numpy as np
import pytest

import oracle as O
from conftest import load_fixture

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


@pytest.fixture(scope="module")
def multi():
    import graphs
    # 6 GMM components (sizes 25..690: one dense-sized, five iterative) + 2 tiny cliques;
    # component offsets ~N(0, I) keep every centroid affinity exp(-d^2) well above
    # underflow, so the meta-embedding's centroid graph is connected and its
    # eigenvectors are unique
    ip, ix, w, x, sizes = graphs.components_graph(6, 20, 800, dim=10, seed=1, device="cpu",
                                                  tiny=(2, 3), spread=1.0)
    return ip, ix, w, x, sizes


def _run(ctx, g, batch, tol=1e-8):
    ip, ix, w, x = g[:4]
    cfg = si.SpectralInitConfig(tol=tol, batch_components=batch)
    return si.spectral_init((ip, ix, w), 2, 42, data=x, config=cfg, ctx=ctx, return_report=True)


def test_batched_components_match_per_component_loop(ctx, multi):
    print("case batched vs loop, sizes", multi[4], flush=True)
    out_b, rep_b = _run(ctx, multi, True)
    out_l, rep_l = _run(ctx, multi, False)
    assert out_b.shape == out_l.shape == (multi[0].size - 1, 2)
    assert rep_b.max_residual <= 1e-8 and rep_l.max_residual <= 1e-8
    # both solve every component to 1e-8; same signs, scale and noise stream
    np.testing.assert_allclose(out_b, out_l, atol=1e-4)


def test_batched_components_match_oracle(ctx, multi):
    print("case batched vs oracle pipeline", flush=True)
    ip, ix, w, x = multi[:4]
    out, _ = _run(ctx, multi, True, tol=1e-6)
    st, ref = O.spectral_init(ip, ix, w, 2, 42, raw_data=x, tol=1e-6)
    assert st == 0
    np.testing.assert_allclose(out, ref, atol=2e-3)


def test_batched_graph_replays_match_eager_rounds(ctx, multi):
    print("case batched hipGraph replays vs eager", flush=True)
    si.batch_graphs(False)
    try:
        out_e, _ = _run(ctx, multi, True)
    finally:
        si.batch_graphs(True)
    for rep in range(3):  # fresh capture per solve, several replays each
        out_g, rep_g = _run(ctx, multi, True)
        cap, launches = si.batch_graphs()
        # at most three graphs (the first FP32 launch, later FP32 launches,
        # the FP64 rounds); each launch replays several rounds planned on the
        # device, so there are fewer launches than rounds
        assert 1 <= cap <= 3 and launches >= cap
        assert rep_g.iterations > launches, (rep_g.iterations, launches)
        np.testing.assert_array_equal(out_g, out_e)
        torch.cuda.synchronize()


def test_batched_reference_disconnected_fixture(ctx):
    print("case batched disconnected_200 fixture", flush=True)
    z = load_fixture("disconnected_200")
    ip, ix, w = z["w_indptr"], z["w_indices"], z["w_data"]
    cfg = si.SpectralInitConfig(tol=1e-6)
    out = si.spectral_init((ip, ix, w), 2, 42, data=z["raw_X"], config=cfg, ctx=ctx)
    st, ref = O.spectral_init(ip, ix, w, 2, 42, raw_data=z["raw_X"], tol=1e-6)
    assert st == 0
    np.testing.assert_allclose(out, ref, atol=2e-3)
