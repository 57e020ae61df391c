"""spectral_init's output stage on the device against its host form.

The stage after the eigensolve -- select_eigenvectors and normalize_signs
(src/selection.rs:11-29, src/scaling.rs:62-74), the placement of every
connected component (src/multi_component.rs:261-312) and scale_and_add_noise
(src/scaling.rs:78-89) -- runs on the device by default; the host form of the
same steps stays reachable through a test seam.  Both must give the same
float32 output bit for bit: one connected graph below and above the
locality-order threshold (host and device component labels), disconnected
graphs with the orthogonal and the spectral meta-embedding, tiny components
placed by the meta-embedding alone, and 2 and 3 embedding dimensions.
"""
import numpy as np
import pytest

from conftest import load_fixture
from helpers import random_symmetric_graph

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


def _fixture(name):
    z = load_fixture(name)
    x = z["raw_X"] if "raw_X" in z else None
    return z["w_indptr"], z["w_indices"], z["w_data"], x


def _grid(a, b):
    # a x b grid graph with unit weights
    import scipy.sparse as sp
    idx = np.arange(a * b).reshape(a, b)
    r = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
    c = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
    m = sp.coo_matrix((np.ones(r.size, np.float32), (r, c)), shape=(a * b, a * b)).tocsr()
    m = (m + m.T).tocsr()
    m.sort_indices()
    return m.indptr.astype(np.int64), m.indices.astype(np.int32), m.data.astype(np.float32), None


def _components(lo, hi, n_comp, seed, tiny=()):
    import graphs
    ip, ix, w, x, _ = graphs.components_graph(n_comp, lo, hi, dim=10, seed=seed, device="cpu",
                                              tiny=tiny, spread=1.0)
    return ip, ix, w, x


CASES = {
    "blobs_connected_2000": lambda: _fixture("blobs_connected_2000"),
    "random_6000": lambda: (*random_symmetric_graph(6000, 12, 3), None),
    "grid_70x71": lambda: _grid(70, 71),
    "disconnected_200": lambda: _fixture("disconnected_200"),
    "components_small_tiny": lambda: _components(20, 800, 6, 1, tiny=(2, 3)),
    "components_large": lambda: _components(800, 4000, 5, 2, tiny=(3,)),
}


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("name", sorted(CASES))
def test_device_output_stage_is_the_host_stage(ctx, name, d):
    print(f"case output stage {name} d={d}", flush=True)
    ip, ix, w, x = CASES[name]()
    cfg = si.SpectralInitConfig(tol=1e-6)
    si.set_output_stage(False)
    try:
        out_h = si.spectral_init((ip, ix, w), d, 42, data=x, config=cfg, ctx=ctx)
    finally:
        si.set_output_stage(True)
    out_d = si.spectral_init((ip, ix, w), d, 42, data=x, config=cfg, ctx=ctx)
    assert out_d.dtype == np.float32 and out_d.shape == (ip.size - 1, d)
    np.testing.assert_array_equal(out_d, out_h)
