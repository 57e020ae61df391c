"""Pins the oracle's warm-start network (oracle/oracle_mpnn.c) independently
of its own code: central finite differences for every gradient, the loss's
invariances, and the features' defining properties."""
import numpy as np
import pytest

import oracle as O
from helpers import random_symmetric_graph


@pytest.fixture(scope="module")
def small():
    ip, ix, w = random_symmetric_graph(60, 5, 4)
    adj, deg = O.adjacency_from_graph(ip, ix, w)
    f = O.mpnn_features(adj, deg, 11)
    rng = np.random.default_rng(0)
    p = rng.standard_normal(O.mpnn_param_count()) * 0.15
    v, _ = np.linalg.qr(rng.standard_normal((60, 2)))
    return adj, deg, f, p, v


def _loss(adj, f, p, v):
    x, _, _ = O.mpnn_forward(adj, f, p)
    return O.mpnn_loss(x, v, grad=False)


def test_param_gradient_matches_central_differences(small):
    adj, _, f, p, v = small
    x, h, av = O.mpnn_forward(adj, f, p)
    _, dx = O.mpnn_loss(x, v)
    g = O.mpnn_backward(adj, f, p, h, av, dx)
    rng = np.random.default_rng(1)
    # every parameter block: Ws1, Wn1, b1, Ws2, Wn2, b2, Ws3, Wn3, b3, Wo
    sizes = [16 * 64, 16 * 64, 64, 64 * 64, 64 * 64, 64, 64 * 64, 64 * 64, 64, 64 * 8]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    idx = np.concatenate([rng.integers(offs[b], offs[b + 1], 4) for b in range(len(sizes))])
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (_loss(adj, f, pp, v) - _loss(adj, f, pm, v)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (i, fd, g[i])


def test_planted_dropped_term_is_caught(small):
    """A backward pass that drops the neighbour term of layer 2 fails the check."""
    adj, _, f, p, v = small
    x, h, av = O.mpnn_forward(adj, f, p)
    _, dx = O.mpnn_loss(x, v)
    g = O.mpnn_backward(adj, f, p, h, av, dx)
    off = 16 * 64 * 2 + 64 + 64 * 64          # start of Wn2
    bad = g.copy()
    bad[off:off + 64 * 64] = 0.0
    i = off + int(np.argmax(np.abs(g[off:off + 64 * 64])))
    eps = 1e-6
    pp, pm = p.copy(), p.copy()
    pp[i] += eps
    pm[i] -= eps
    fd = (_loss(adj, f, pp, v) - _loss(adj, f, pm, v)) / (2 * eps)
    assert abs(fd - g[i]) < 1e-6 * max(1, abs(fd)) + 1e-8
    assert abs(fd - bad[i]) > 1e-4


def test_loss_gradient_matches_central_differences(small):
    adj, _, f, p, v = small
    x, _, _ = O.mpnn_forward(adj, f, p)
    loss, dx = O.mpnn_loss(x, v)
    eps = 1e-7
    for i, j in [(0, 0), (7, 1), (33, 5), (59, 7), (12, 3)]:
        xp, xm = x.copy(), x.copy()
        xp[i, j] += eps
        xm[i, j] -= eps
        fd = (O.mpnn_loss(xp, v, grad=False) - O.mpnn_loss(xm, v, grad=False)) / (2 * eps)
        assert abs(fd - dx[i, j]) <= 1e-6 * max(1.0, abs(fd))


def test_loss_invariances(small):
    _, _, _, _, v = small
    rng = np.random.default_rng(2)
    x = rng.standard_normal((v.shape[0], 8))
    base = O.mpnn_loss(x, v, grad=False)
    c, s = np.cos(0.7), np.sin(0.7)
    assert np.isclose(O.mpnn_loss(x, v @ np.array([[c, -s], [s, c]]), grad=False), base)
    assert np.isclose(O.mpnn_loss(x, -v, grad=False), base)
    assert np.isclose(O.mpnn_loss(3.5 * x, v, grad=False), base)
    q, _ = np.linalg.qr(rng.standard_normal((8, 8)))
    assert np.isclose(O.mpnn_loss(x @ q, v, grad=False), base)   # any orthonormal mix of X


def test_loss_closed_forms(small):
    """tr(V^T V) - tr(B^T (G + delta I)^-1 B) on bases where it is known."""
    _, _, _, _, v = small
    n = v.shape[0]
    rng = np.random.default_rng(3)
    w = rng.standard_normal((n, 7))
    w -= v @ (v.T @ w)
    w, _ = np.linalg.qr(w)                     # orthonormal, orthogonal to V
    # X = [V | W_6]: G = I, delta = 1e-3, loss = 2 - 2 / (1 + 1e-3)
    x = np.column_stack([v, w[:, :6]])
    assert np.isclose(O.mpnn_loss(x, v, grad=False), 2.0 - 2.0 / 1.001, rtol=1e-12)
    # X orthogonal to V: nothing covered
    x = np.column_stack([w[:, :6], w[:, :2] * 0.5])
    assert np.isclose(O.mpnn_loss(x, v, grad=False), 2.0, rtol=1e-12)
    # one eigenvector covered, X = [v_1 | W_7] orthonormal: loss = 2 - 1 / (1 + 1e-3)
    x = np.column_stack([v[:, 0], w])
    assert np.isclose(O.mpnn_loss(x, v, grad=False), 2.0 - 1.0 / 1.001, rtol=1e-12)
    # scaled to 3x: unchanged (delta scales with tr(G))
    assert np.isclose(O.mpnn_loss(3.0 * x, v, grad=False), 2.0 - 1.0 / 1.001, rtol=1e-12)


def test_features_definition(small):
    adj, deg, f, _, _ = small
    ip, ix, a = adj
    sq = np.sqrt(deg)
    assert np.allclose(f[:, 0], np.log1p(deg))
    assert np.allclose(f[:, 1], np.log1p(np.diff(ip)))
    assert np.isclose(f[:, 14].max(), 1.0) and np.allclose(f[:, 15], 1.0)
    i = 5
    row = a[ip[i]:ip[i + 1]]
    assert np.allclose(f[i, 2:6], [row.mean(), row.max(), row.min(), row.std()])
    for t in range(8):
        pt = f[:, 6 + t]
        assert np.isclose(np.sqrt(np.mean(pt ** 2)), 1.0)
        # the trivial direction stays projected out (A_hat D^1/2 1 = D^1/2 1 up to
        # the FP32 rounding of the PythonCompat degrees)
        assert abs(pt @ sq) / (np.linalg.norm(pt) * np.linalg.norm(sq)) < 1e-6
    # the probes are A_hat^t applied to the deflated draws, t = 1, 2, 3, 4, 6, 8, 12, 16
    import scipy.sparse as sp
    n = ip.size - 1
    A = sp.csr_matrix((a, ix, ip), shape=(n, n))
    p = O.normal_block(11, 9, n)
    p = p - (p @ sq) / (sq @ sq) * sq
    done = 0
    for col, t in enumerate([1, 2, 3, 4, 6, 8, 12, 16]):
        while done < t:
            p = A @ p
            done += 1
        assert np.allclose(f[:, 6 + col], p / np.sqrt(np.mean(p ** 2)))
