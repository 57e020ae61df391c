"""Pins of the oracle's residual (the escalation gate) and of its RNG.

The residual ||L v - lambda v|| / max(||v||, 1e-300) decides every
escalation level and the eigensolver contract's "residual <= 1e-6"
(src/metrics.rs:57-92; per-column form in src/solvers/lobpcg.rs:193-223).
It is pinned by closed forms that a plausible slip breaks:

* path P3, combinatorial Laplacian [[1,-1,0],[-1,2,-1],[0,-1,1]], v = e1,
  lambda = 0: L v = (1, -1, 0), residual sqrt(2) (metrics.rs:401-408, 444-452
  use the same matrix and vector);
* the same with lambda = 1: L v - v = (0, -1, 0), residual 1 (a flipped sign
  of the lambda term gives sqrt(5));
* P3's normalized Laplacian, v = e1, lambda = 0: (1, -1/sqrt 2, 0), residual
  sqrt(1.5);
* scale invariance (a residual divided by ||v||^2 is not);
* the worst column in either position (metrics.rs:456-476);
* NaN in any column or eigenvalue propagates (metrics.rs:80-91);
* agreement with numpy's dense ||L V - V diag(lambda)|| / ||V|| on random
  blocks.

planted_* tests restate the residual with one mistake each (returns 0,
divides by ||v||^2, reads only the first column, ignores NaN, flips the sign
of lambda) and show that the pins above reject every one of them.

The RNG: Philox4x32-10 against the known-answer vectors published with
Random123 (Salmon et al., SC'11; kat_vectors), and the Box-Muller normals by a
Kolmogorov-Smirnov test against the standard normal CDF.
"""
import math

import numpy as np
import pytest
import scipy.stats

import oracle as O
from helpers import path_graph, random_symmetric_graph


def _csr(dense):
    dense = np.asarray(dense, float)
    n = dense.shape[0]
    ip, ix, d = [0], [], []
    for i in range(n):
        for j in range(n):
            if dense[i, j] != 0.0:
                ix.append(j)
                d.append(dense[i, j])
        ip.append(len(ix))
    return (np.array(ip, np.int64), np.array(ix, np.int32), np.array(d, np.float64))


P3_COMB = _csr([[1, -1, 0], [-1, 2, -1], [0, -1, 1]])


def _p3_normalized():
    ip, ix, w = path_graph(3)
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    return L


def _random_case(seed, n=300, k=5):
    ip, ix, w = random_symmetric_graph(n, 8, seed)
    L, _, _ = O.laplacian_from_graph(ip, ix, w)
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, k)) * rng.uniform(0.01, 100.0, k)[None, :]
    lam = rng.uniform(0.0, 2.0, k)
    return L, lam, v


def _dense(L):
    ip, ix, d = L
    n = ip.size - 1
    out = np.zeros((n, n))
    for i in range(n):
        out[i, ix[ip[i]:ip[i + 1]]] += d[ip[i]:ip[i + 1]]
    return out


def residual_pin_failures(fn):
    """Names of the closed-form pins that fn(L, evals, evecs) -> max residual
    fails; an empty list for a correct max_eigenpair_residual."""
    fails = []
    e1 = np.array([[1.0], [0.0], [0.0]])
    s = 1.0 / math.sqrt(3.0)
    triv = np.array([[s], [s], [s]])

    def check(name, got, want, tol=1e-14):
        if not (abs(got - want) <= tol * max(1.0, abs(want))):
            fails.append(name)

    check("p3_e1", fn(P3_COMB, np.array([0.0]), e1), math.sqrt(2.0))
    check("p3_e1_lambda1", fn(P3_COMB, np.array([1.0]), e1), 1.0)
    check("p3_normalized_e1", fn(_p3_normalized(), np.array([0.0]), e1), math.sqrt(1.5))
    if not fn(P3_COMB, np.array([0.0]), triv) < 1e-10:
        fails.append("p3_trivial")
    for c in (1e-3, 7.5):
        check(f"scale_{c}", fn(P3_COMB, np.array([0.0]), c * e1), math.sqrt(2.0))
    both = np.hstack([triv, e1])
    check("worst_second", fn(P3_COMB, np.array([0.0, 0.0]), both), math.sqrt(2.0))
    check("worst_first", fn(P3_COMB, np.array([0.0, 0.0]), both[:, ::-1].copy()), math.sqrt(2.0))
    nanv = np.hstack([e1, np.array([[np.nan], [0.0], [0.0]])])
    if not math.isnan(fn(P3_COMB, np.array([0.0, 0.0]), nanv)):
        fails.append("nan_column_last")
    if not math.isnan(fn(P3_COMB, np.array([0.0, 0.0]), nanv[:, ::-1].copy())):
        fails.append("nan_column_first")
    if not math.isnan(fn(P3_COMB, np.array([0.0, np.nan]), np.hstack([e1, e1]))):
        fails.append("nan_eigenvalue")
    for seed in (1, 2):
        L, lam, v = _random_case(seed)
        want = np.max(np.linalg.norm(_dense(L) @ v - v * lam[None, :], axis=0)
                      / np.linalg.norm(v, axis=0))
        check(f"random_{seed}", fn(L, lam, v), want, tol=1e-12)
    return fails


def test_oracle_residual_passes_every_pin():
    assert residual_pin_failures(O.max_eigenpair_residual) == []


def test_oracle_p3_closed_forms_exact():
    e1 = np.array([[1.0], [0.0], [0.0]])
    assert O.max_eigenpair_residual(P3_COMB, np.array([0.0]), e1) == math.sqrt(2.0)
    assert O.max_eigenpair_residual(_p3_normalized(), np.array([0.0]), e1) == pytest.approx(
        math.sqrt(1.5), rel=1e-15)
    # the reference's own bars (metrics.rs:444-476): >= RSVD_QUALITY_THRESHOLD
    assert O.max_eigenpair_residual(P3_COMB, np.array([0.0]), e1) >= 1e-2


def test_oracle_zero_vector_uses_the_norm_floor():
    # ||v|| = 0: the difference is 0 as well, 0 / 1e-300 = 0 (metrics.rs:66)
    z = np.zeros((3, 1))
    assert O.max_eigenpair_residual(P3_COMB, np.array([0.3]), z) == 0.0
    # a tiny v keeps its scale invariance down to the floor
    e1 = np.array([[1e-150], [0.0], [0.0]])
    assert O.max_eigenpair_residual(P3_COMB, np.array([0.0]), e1) == pytest.approx(math.sqrt(2.0))


def test_per_vector_residuals_match_numpy():
    L, lam, v = _random_case(9)
    res = O.per_vector_residuals(L, lam, v)
    want = np.linalg.norm(_dense(L) @ v - v * lam[None, :], axis=0) / np.linalg.norm(v, axis=0)
    np.testing.assert_allclose(res, want, rtol=1e-12)
    assert O.max_eigenpair_residual(L, lam, v) == pytest.approx(want.max(), rel=1e-12)


# ------------------------------------------------------------------ planted mistakes
def _planted(kind):
    def fn(L, evals, evecs):
        d = _dense(L)
        v = np.asarray(evecs, float)
        lam = np.asarray(evals, float)
        cols = range(v.shape[1])
        if kind == "first_column_only":
            cols = [0]
        out = 0.0
        for c in cols:
            r = d @ v[:, c] + (1 if kind == "sign" else -1) * lam[c] * v[:, c]
            vn = np.linalg.norm(v[:, c])
            if kind == "norm_squared":
                vn = vn * vn
            val = np.linalg.norm(r) / max(vn, 1e-300)
            if kind == "returns_zero":
                val = 0.0
            if kind == "nan_ignored":
                out = max(out, val) if not math.isnan(val) else out
            else:
                out = math.nan if (math.isnan(out) or math.isnan(val)) else max(out, val)
        return out
    return fn


@pytest.mark.parametrize("kind", ["returns_zero", "norm_squared", "first_column_only",
                                  "nan_ignored", "sign"])
def test_planted_residual_mistakes_are_caught(kind):
    assert residual_pin_failures(_planted(kind)), kind


def test_planted_restatement_without_mistake_passes():
    # the same restatement with no mistake passes: the pins test behaviour
    assert residual_pin_failures(_planted("none")) == []


# ------------------------------------------------------------------ RNG
# Random123 known-answer vectors for philox4x32-10 (counter, key, output)
PHILOX_KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    got = O.philox4x32_10(ctr, key[0], key[1])
    assert [int(v) for v in got] == list(want)


def test_normals_follow_the_standard_normal_law():
    x = O.normal_block(42, 1, 200000)
    ks = scipy.stats.kstest(x, "norm")
    assert ks.pvalue > 1e-3, ks
    # other streams and seeds are independent draws of the same law
    y = O.normal_block(42, 2, 200000)
    assert scipy.stats.kstest(y, "norm").pvalue > 1e-3
    assert abs(np.corrcoef(x, y)[0, 1]) < 0.01
    assert np.isfinite(x).all() and np.abs(x).max() < 7.0


def test_fp32_pair_stream_from_the_pinned_philox():
    # the FP32 start-block stream: pair p takes the first two Philox words of
    # counter (p, 0, stream, 0x9a1f) under the seed key (the Philox itself is
    # pinned by the known answers above) as 24-bit uniforms u1 in (0, 1] and
    # u2 in [0, 1); its two normals have squared length -2 ln u1 and angle
    # 2 pi u2 (the Box-Muller identities), each to float32 rounding
    seed, stream, npairs = 42, 1, 64
    z = O.normal_pairs_f32(seed, stream, npairs).astype(np.float64)
    for p in range(npairs):
        c = O.philox4x32_10((p, 0, stream, 0x9A1F), seed & 0xFFFFFFFF, seed >> 32)
        u1 = ((int(c[0]) >> 8) + 1) / 2.0 ** 24
        u2 = (int(c[1]) >> 8) / 2.0 ** 24
        r2 = z[p, 0] ** 2 + z[p, 1] ** 2
        assert abs(r2 - (-2.0 * np.log(u1))) <= 1e-5 * max(r2, 1.0), (p, r2, u1)
        ang = np.arctan2(z[p, 1], z[p, 0]) % (2 * np.pi)
        d = abs(ang - 2 * np.pi * u2)
        assert min(d, 2 * np.pi - d) < 1e-5 * max(1.0, 1.0 / np.sqrt(r2)), (p, ang, u2)


def test_fp32_pair_stream_follows_the_standard_normal_law():
    z = O.normal_pairs_f32(42, 1, 100000)
    assert z.dtype == np.float32
    for col in (z[:, 0], z[:, 1]):
        assert scipy.stats.kstest(col.astype(np.float64), "norm").pvalue > 1e-3
    # the two outputs of a pair are uncorrelated, and so are two streams
    assert abs(np.corrcoef(z[:, 0], z[:, 1])[0, 1]) < 0.01
    w = O.normal_pairs_f32(42, 2, 100000)
    assert abs(np.corrcoef(z[:, 0], w[:, 0])[0, 1]) < 0.01
