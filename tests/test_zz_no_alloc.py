"""si_max_eigenpair_residual_noalloc against si_max_eigenpair_residual.

The allocation-free entry point keeps its partial sums in a static device
array shared by every call on the GPU.  Two calls in a row, the second on a
smaller input, must each equal the allocating entry point's result (FP64,
1e-12 relative: the two sum in different orders), so nothing of the first
call's scratch leaks into the second.  Closed forms, NaN propagation and the
size limit are checked too.
"""
import math

import numpy as np
import pytest

import oracle as O
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


def _case(ctx, n, k, seed):
    ip, ix, w = random_symmetric_graph(n, 10, seed)
    lap = si.Laplacian(ctx, ip, ix, w)  # n >= 4096 takes the locality order
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, k)) * rng.uniform(1e-2, 1e2, k)[None, :]
    lam = rng.uniform(0.0, 2.0, k)
    return lap, L, lam, v


def test_noalloc_matches_allocating_entry_two_calls(ctx):
    print("case noalloc n=6000 k=13 then n=3000 k=3", flush=True)
    big = _case(ctx, 6000, 13, 11)
    small = _case(ctx, 3000, 3, 12)
    for lap, L, lam, v in (big, small):
        vd = _dev(v)
        want = si.max_eigenpair_residual(lap, lam, vd)
        got = si.max_eigenpair_residual_noalloc(lap, lam, vd)
        assert got == pytest.approx(want, rel=1e-12)
        assert got == pytest.approx(O.max_eigenpair_residual(L, lam, v), rel=1e-12)
    # and again on the larger input after the smaller one
    lap, L, lam, v = big
    vd = _dev(v)
    assert si.max_eigenpair_residual_noalloc(lap, lam, vd) == pytest.approx(
        si.max_eigenpair_residual(lap, lam, vd), rel=1e-12)


def test_noalloc_closed_form_and_nan(ctx):
    ip, ix, w = path_graph(3)
    lap = si.Laplacian(ctx, ip, ix, w)
    e1 = np.array([[1.0], [0.0], [0.0]])
    assert si.max_eigenpair_residual_noalloc(lap, np.array([0.0]), _dev(e1)) == pytest.approx(
        math.sqrt(1.5), rel=1e-15)
    nanv = np.hstack([e1, np.array([[np.nan], [0.0], [0.0]])])
    assert math.isnan(si.max_eigenpair_residual_noalloc(lap, np.zeros(2), _dev(nanv)))
    assert math.isnan(si.max_eigenpair_residual_noalloc(lap, np.array([0.0, np.nan]),
                                                        _dev(np.hstack([e1, e1]))))


def test_noalloc_64_columns_and_limit(ctx):
    lap, L, lam, v = _case(ctx, 2000, 64, 5)
    vd = _dev(v)
    assert si.max_eigenpair_residual_noalloc(lap, lam, vd) == pytest.approx(
        si.max_eigenpair_residual(lap, lam, vd), rel=1e-12)
    wide = _dev(np.ones((2000, 65)))
    with pytest.raises(ValueError, match="outside"):
        si.max_eigenpair_residual_noalloc(lap, np.zeros(65), wide)
