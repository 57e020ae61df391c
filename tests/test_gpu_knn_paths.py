"""GPU checks of the two exact k-NN paths of si_knn_exact (float32 and float64
points)
(spectral-init_amd/csrc/k_graph.hip): the FP32 MFMA filter with the certified
FP64 re-rank, and the direct FP64 kernel.  Every path must return the direct
kernel's result bit for bit; the direct kernel itself is pinned against the
oracle and the reference fixtures (test_gpu_graph.py), and the cases below that
fit the oracle are checked against it too.

Stage 1 runs on the bf16 matrix cores (three bf16 products per FP32 product)
and falls back to the FP32 matrix cores when that leaves more than 1 % of the
queries uncertified (or the norms are too large for it); mode 3 forces the FP32
stage.  The certificate sends a query to the direct kernel when one of its kept
candidate lists might have dropped a true neighbour: heavy exact ties,
coordinates whose squares underflow in float32, and coordinates whose squares
overflow it (those skip the filter altogether) are the inputs that exercise
it."""
import numpy as np
import pytest

import oracle as O
import spectral_mi355x as si

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    c = si.Context(0)
    yield c
    si.knn_path(0)
    c.close()


def run(ctx, x, k, mode):
    si.knn_path(mode)
    try:
        idx, dist = si.knn_exact(ctx, x, k)
        return idx, dist, si.knn_path(-1)
    finally:
        si.knn_path(0)


def gmm(n, d, clusters, seed, scale=3.0):
    rng = np.random.default_rng(seed)
    centres = rng.standard_normal((clusters, d)) * scale
    lab = rng.integers(0, clusters, n)
    return (centres[lab] + rng.standard_normal((n, d))).astype(np.float32)


@pytest.mark.parametrize("d", [30, 50])
def test_filter_equals_direct_kernel_100k(ctx, d):
    print(f"case knn paths 100k d={d}", flush=True)
    x = gmm(100_000, d, 20, d)
    x[17] = x[99_000]          # an exact duplicate pair
    x[500:520] = x[500]        # twenty copies: ties broken by index
    i1, d1, _ = run(ctx, x, 15, 1)
    i0, d0, direct = run(ctx, x, 15, 0)
    np.testing.assert_array_equal(i0, i1)
    np.testing.assert_array_equal(d0.view(np.uint32), d1.view(np.uint32))
    # the certificate holds for nearly every query of well-spread points
    assert direct < 0.01 * x.shape[0], direct
    i3, d3, direct3 = run(ctx, x, 15, 3)     # the FP32 stage 1
    np.testing.assert_array_equal(i3, i1)
    np.testing.assert_array_equal(d3.view(np.uint32), d1.view(np.uint32))
    assert direct3 < 0.01 * x.shape[0], direct3
    assert i0[17, 0] == 99_000 and i0[99_000, 0] == 17
    np.testing.assert_array_equal(i0[500, :15], np.arange(501, 516))


def test_forced_direct_requeries_every_query(ctx):
    print("case knn paths forced direct", flush=True)
    x = gmm(20_000, 17, 8, 5)
    i1, d1, _ = run(ctx, x, 10, 1)
    i2, d2, direct = run(ctx, x, 10, 2)
    assert direct == x.shape[0]
    np.testing.assert_array_equal(i2, i1)
    np.testing.assert_array_equal(d2.view(np.uint32), d1.view(np.uint32))


@pytest.mark.parametrize("k", [1, 8, 16, 17, 30, 32, 33])
def test_k_boundary_matches_oracle(ctx, k):
    print(f"case knn paths k={k}", flush=True)
    x = gmm(3000, 13, 6, 100 + k)
    idx, dist, _ = run(ctx, x, k, 0)
    oi, od = O.knn_exact(x.astype(np.float64), k)
    np.testing.assert_array_equal(idx, oi)
    np.testing.assert_allclose(dist, od, rtol=1.2e-7, atol=0)


def test_heavy_ties_fall_back_and_match_oracle(ctx):
    print("case knn paths heavy ties", flush=True)
    # 100 consecutive copies of each of 40 points: 99 candidates at distance 0,
    # more than one kept list of 16 can hold, in every query's own range
    base = gmm(40, 7, 3, 9)
    x = np.repeat(base, 100, axis=0)
    idx, dist, direct = run(ctx, x, 15, 0)
    oi, od = O.knn_exact(x.astype(np.float64), 15)
    np.testing.assert_array_equal(idx, oi)
    np.testing.assert_allclose(dist, od, rtol=1.2e-7, atol=0)
    assert direct > 0


@pytest.mark.parametrize("scale", [1e-25, 1e20])
def test_extreme_scales_match_oracle(ctx, scale):
    print(f"case knn paths scale={scale}", flush=True)
    x = (gmm(2000, 9, 4, 3).astype(np.float64) * scale).astype(np.float32)
    idx, dist, _ = run(ctx, x, 15, 0)
    oi, od = O.knn_exact(x.astype(np.float64), 15)
    np.testing.assert_array_equal(idx, oi)
    np.testing.assert_allclose(dist, od, rtol=1.2e-7, atol=0)


def test_outlier_points_match_oracle(ctx):
    """A few points far from the rest widen every query's error bound."""
    print("case knn paths outliers", flush=True)
    x = gmm(4000, 30, 10, 21)
    x[:5] *= 1e4
    idx, dist, _ = run(ctx, x, 15, 0)
    oi, od = O.knn_exact(x.astype(np.float64), 15)
    np.testing.assert_array_equal(idx, oi)
    np.testing.assert_allclose(dist, od, rtol=1.2e-7, atol=0)


@pytest.mark.parametrize("d", [8, 32])
def test_far_clusters_take_the_fp32_stage_and_match_oracle(ctx, d):
    """Two tight clusters 2000 units apart: the bf16 parts cannot resolve the
    neighbours against |q'|^2 ~ 1e6, so the FP32 stage (and the direct kernel
    for what it leaves) must produce the result.  At d = 32 the two stages use
    different queries per workgroup."""
    print(f"case knn paths far clusters d={d}", flush=True)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((3000, d))
    x[:1500, 0] += 1000.0
    x[1500:, 0] -= 1000.0
    x = x.astype(np.float32)
    oi, od = O.knn_exact(x.astype(np.float64), 15)
    for mode in (0, 3):
        idx, dist, _ = run(ctx, x, 15, mode)
        np.testing.assert_array_equal(idx, oi)
        np.testing.assert_allclose(dist, od, rtol=1.2e-7, atol=0)


@pytest.mark.parametrize("d", [1, 2, 30, 31, 62, 63, 64])
def test_widths_across_k_blocks_match_oracle(ctx, d):
    """One, two and three 32-slot k-blocks of the bf16 stage (d + 2 norm slots),
    at and around each block boundary."""
    print(f"case knn paths width d={d}", flush=True)
    x = gmm(2500, d, 5, 300 + d)
    x[11] = x[2222]
    idx, dist, _ = run(ctx, x, 15, 0)
    oi, od = O.knn_exact(x.astype(np.float64), 15)
    np.testing.assert_array_equal(idx, oi)
    np.testing.assert_allclose(dist, od, rtol=1.2e-7, atol=0)


def test_float64_points_filter_equals_direct_kernel(ctx):
    """float64 points (the reference fixtures' type) take the same filter: the
    centred coordinates are rounded to float32 for stage 1, and the re-rank uses
    the float64 coordinates, exactly as the direct kernel does."""
    print("case knn paths float64 100k", flush=True)
    x = gmm(100_000, 30, 20, 77).astype(np.float64)
    x += np.random.default_rng(5).standard_normal(x.shape) * 1e-9  # not float32 values
    x[3] = x[77_777]
    i1, d1, _ = run(ctx, x, 15, 1)
    i0, d0, direct = run(ctx, x, 15, 0)
    np.testing.assert_array_equal(i0, i1)
    np.testing.assert_array_equal(d0.view(np.uint32), d1.view(np.uint32))
    assert direct < 0.01 * x.shape[0], direct
    assert i0[3, 0] == 77_777


@pytest.mark.parametrize("d", [1, 2, 30])
def test_float64_widths_match_oracle(ctx, d):
    print(f"case knn paths float64 d={d}", flush=True)
    x = gmm(2500, d, 5, 900 + d).astype(np.float64) * 1.0000001
    idx, dist, _ = run(ctx, x, 15, 0)
    oi, od = O.knn_exact(x, 15)
    np.testing.assert_array_equal(idx, oi)
    np.testing.assert_allclose(dist, od, rtol=1.2e-7, atol=0)
