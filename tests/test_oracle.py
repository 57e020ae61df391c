"""Pin the FP64 CPU oracle before it is trusted as the checker.

Each check compares the oracle with something it does not share code with:
the reference's golden fixtures (Python UMAP / scipy outputs, Rust-produced
eigenvectors), closed forms (ring spectra, Chebyshev polynomials, rotated
bases) or library routines (numpy / scipy eigh, dense products).
"""
import numpy as np
import numpy.polynomial.chebyshev as npcheb
import pytest
import scipy.linalg
import scipy.sparse as sp

import oracle as O
from conftest import DATASETS, load_fixture
from helpers import (csr_to_dense, max_subspace_angle, path_graph, principal_angles,
                     random_symmetric_graph, ring_eigenvalues, ring_graph)


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


def _lap(z):
    deg, sq = O.compute_degrees(z["w_indptr"], z["w_indices"], z["w_data"], 0)
    isd = np.where(sq > 0, 1.0 / np.where(sq > 0, sq, 1.0), 0.0)
    return O.build_laplacian(z["w_indptr"], z["w_indices"], z["w_data"], isd), deg, sq


# --------------------------------------------------------------- a1 / a2
@pytest.mark.parametrize("name", DATASETS)
def test_degrees_match_fixture_python_compat(name):
    # test_comp_a_degrees.rs:41-47 -- PythonCompat degrees to 1e-15 relative
    z = load_fixture(name)
    deg, sq = O.compute_degrees(z["w_indptr"], z["w_indices"], z["w_data"], 0)
    np.testing.assert_allclose(deg, z["degrees"], rtol=1e-15, atol=0)
    np.testing.assert_allclose(sq, z["sqrt_deg"], rtol=1e-15, atol=0)


@pytest.mark.parametrize("name", DATASETS)
def test_degrees_rust_native_close(name):
    # test_comp_a_degrees.rs:68 -- RustNative within 3e-7 relative
    z = load_fixture(name)
    deg, _ = O.compute_degrees(z["w_indptr"], z["w_indices"], z["w_data"], 1)
    ref = z["degrees"]
    np.testing.assert_allclose(deg, ref, rtol=3e-7)
    # and it is the exact f64 row sum
    w = sp.csr_matrix((z["w_data"].astype(np.float64), z["w_indices"], z["w_indptr"]))
    np.testing.assert_allclose(deg, np.asarray(w.sum(axis=1)).ravel(), rtol=1e-14)


@pytest.mark.parametrize("name", DATASETS)
def test_laplacian_matches_fixture(name):
    # test_comp_b_laplacian.rs -- structure identical, values to 1e-14
    z = load_fixture(name)
    (ip, ix, d), _, _ = _lap(z)
    np.testing.assert_array_equal(ip, z["l_indptr"])
    np.testing.assert_array_equal(ix, z["l_indices"])
    np.testing.assert_allclose(d, z["l_data"], rtol=0, atol=1e-14)


def test_laplacian_diagonal_for_isolated_node():
    # laplacian.rs:83-85: every diagonal entry is 1, isolated nodes included
    ip = np.array([0, 1, 2, 2], np.int64)
    ix = np.array([1, 0], np.int32)
    w = np.array([0.5, 0.5], np.float32)
    (lip, lix, ld), deg, sq = _lap(dict(w_indptr=ip, w_indices=ix, w_data=w))
    dense = csr_to_dense(lip, lix, ld, 3)
    np.testing.assert_allclose(np.diag(dense), [1, 1, 1])
    np.testing.assert_allclose(dense[0, 1], -1.0)
    assert deg[2] == 0 and sq[2] == 0


# --------------------------------------------------------------- a3
def test_spmm_matches_dense_product():
    ip, ix, w = random_symmetric_graph(300, 12, 3)
    (lip, lix, ld), _, _ = _lap(dict(w_indptr=ip, w_indices=ix, w_data=w))
    dense = csr_to_dense(lip, lix, ld, 300)
    x = np.random.default_rng(0).standard_normal((300, 5))
    np.testing.assert_allclose(O.spmm((lip, lix, ld), x), dense @ x, rtol=0, atol=1e-13)
    # block = column-by-column (operator.rs k=1 path equals k>1 path)
    cols = np.stack([O.spmm((lip, lix, ld), x[:, c]) for c in range(5)], axis=1)
    np.testing.assert_array_equal(cols, O.spmm((lip, lix, ld), x))


# --------------------------------------------------------------- a5
def _diag_csr(vals):
    n = vals.size
    return (np.arange(n + 1, dtype=np.int64), np.arange(n, dtype=np.int32), vals.astype(np.float64))


@pytest.mark.parametrize("degree", [1, 2, 5, 8])
def test_chebyshev_filter_closed_form_on_diagonal(degree):
    # For diagonal L the filter is the scalar T_d((lambda - c)/e) per row.
    lam = np.linspace(0.0, 2.0, 41)
    r = np.random.default_rng(1).standard_normal((41, 3))
    out = O.chebyshev_filter(_diag_csr(lam), r, 0.1, 2.0, degree)
    c, e = 1.05, 0.95
    coef = np.zeros(degree + 1)
    coef[-1] = 1.0
    t = npcheb.chebval((lam - c) / e, coef)
    np.testing.assert_allclose(out, t[:, None] * r, rtol=1e-12, atol=1e-12)


def test_chebyshev_amplifies_low_band():
    # lobpcg.rs:615-642: components below a are amplified >= 10x vs those in [a, b]
    lam = np.array([0.0, 0.01, 0.5, 1.0, 1.9])
    out = O.chebyshev_filter(_diag_csr(lam), np.eye(5), 0.1, 2.0, 8)
    gains = np.abs(np.diag(out))
    assert gains[0] >= 10 * gains[2:].max()


# --------------------------------------------------------------- dense kernels
def test_sym_eig_matches_numpy():
    a = np.random.default_rng(2).standard_normal((24, 24))
    a = a + a.T
    w, v = O.sym_eig(a)
    np.testing.assert_allclose(w, np.linalg.eigvalsh(a), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(a @ v, v * w, atol=1e-11)


def test_gen_sym_eig_matches_scipy():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((15, 15)); a = a + a.T
    b = rng.standard_normal((15, 15)); b = b @ b.T + 15 * np.eye(15)
    w, c = O.gen_sym_eig(a, b)
    np.testing.assert_allclose(w, scipy.linalg.eigh(a, b, eigvals_only=True), rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(c.T @ b @ c, np.eye(15), atol=1e-10)


@pytest.mark.parametrize("name", ["blobs_50", "moons_200", "circles_300", "blobs_connected_200",
                                  "near_dupes_100"])
def test_dense_evd_matches_numpy_and_fixture(name):
    # test_comp_d_dense_evd.rs:7-66: eigenvalues 1e-9, residual 1e-10
    z = load_fixture(name)
    L, _, _ = _lap(z)
    n = L[0].size - 1
    ev, vec = O.dense_evd(L, 3)
    dense = csr_to_dense(*L, n)
    np.testing.assert_allclose(ev, np.linalg.eigvalsh(dense)[:3], atol=1e-12)
    # the fixture generator clips eigenvalues at 0 (generate_fixtures.py:298)
    np.testing.assert_allclose(np.maximum(ev, 0.0), z["eig_values"], atol=1e-9)
    assert O.max_eigenpair_residual(L, ev, vec) < 1e-10


# --------------------------------------------------------------- a8
def test_rayleigh_ritz_recovers_rotated_basis():
    # lobpcg.rs:672-759: RR on an exact eigenbasis mixed by a rotation
    ip, ix, w = random_symmetric_graph(200, 10, 4)
    L, _, _ = _lap(dict(w_indptr=ip, w_indices=ix, w_data=w))
    dense = csr_to_dense(*L, 200)
    lam, vec = np.linalg.eigh(dense)
    q, _ = np.linalg.qr(np.random.default_rng(5).standard_normal((3, 3)))
    ev, x = O.rayleigh_ritz(L, vec[:, :3] @ q)
    np.testing.assert_allclose(ev, lam[:3], atol=1e-12)
    for c in range(3):
        assert abs(abs(x[:, c] @ vec[:, c]) - 1.0) < 1e-10
    res = O.per_vector_residuals(L, ev, x)
    assert res.max() < 1e-10


# --------------------------------------------------------------- a6 / a7 / a10
@pytest.mark.parametrize("name", ["blobs_connected_2000", "blobs_5000", "blobs_500",
                                  "circles_300", "moons_200", "blobs_connected_200"])
def test_lobpcg_solve_matches_scipy_fixture(name):
    # test_comp_f_lobpcg.rs:22-48: eigenvalues vs scipy 1e-6 abs, residual < 1e-5
    z = load_fixture(name)
    L, _, sq = _lap(z)
    r = O.lobpcg_solve(L, 2, 42, sq)
    assert r is not None
    np.testing.assert_allclose(r["evals"], z["eig_values"], atol=1e-6)
    assert O.max_eigenpair_residual(L, r["evals"], r["evecs"]) < 1e-5


def test_lobpcg_matches_rust_produced_vectors():
    # Rust L1-LOBPCG output of blobs_connected_2000 (sign-normalised, x10 scaled,
    # 1e-4 noise): its subspace agrees with ours to the noise floor (~1e-4 rad).
    z = load_fixture("blobs_connected_2000")
    L, _, sq = _lap(z)
    r = O.lobpcg_solve(L, 2, 42, sq)
    rust = np.load("tests/golden/rust_vectors/blobs_connected_2000_scalar.npy").astype(np.float64)
    ang = principal_angles(r["evecs"], rust)
    assert ang[0] < 3e-4, ang


def test_dense_and_lobpcg_agree_on_ring200():
    # test_rq3_threshold.rs:33-90: forced-threshold ring(200) within 0.01 rad,
    # and the Rust-produced dense / LOBPCG blocks span the same subspace.
    ip, ix, w = ring_graph(200)
    L, _, sq = _lap(dict(w_indptr=ip, w_indices=ix, w_data=w))
    dense = O.solve_eigenproblem(L, 2, 42, sq, dense_threshold=2000)
    assert dense["level"] == 0
    np.testing.assert_allclose(dense["evals"], ring_eigenvalues(200, 3), atol=1e-12)
    rust_dense = np.load("tests/golden/rust_vectors/ring200_dense_eigvecs.npy")
    # lambda_1 = lambda_2 (degenerate pair): compare the whole 3-d subspace
    assert max_subspace_angle(dense["evecs"], rust_dense) < 1e-7


@pytest.mark.parametrize("n,tol", [(600, 1e-9), (1200, 1e-9), (3000, 1e-7)])
def test_lobpcg_on_ring_closed_form(n, tol):
    # ring C_n: lambda_j = 1 - cos(2 pi j / n) with degenerate pairs.  At a
    # tight residual tolerance the eigenvalue error is bounded by res^2 / gap,
    # so the closed form must be met to ~1e-12.
    ip, ix, w = ring_graph(n)
    L, _, sq = _lap(dict(w_indptr=ip, w_indices=ix, w_data=w))
    r = O.lobpcg_solve(L, 2, 42, sq, tol=tol)
    assert r is not None and r["converged"]
    np.testing.assert_allclose(r["evals"], ring_eigenvalues(n, 3), atol=5e-12)
    assert O.max_eigenpair_residual(L, r["evals"], r["evecs"]) < tol


def test_chain_on_ring_meets_gate():
    # with the reference tolerance (1e-5) the chain's answer passes its gate
    ip, ix, w = ring_graph(2500)
    L, _, sq = _lap(dict(w_indptr=ip, w_indices=ix, w_data=w))
    r = O.solve_eigenproblem(L, 2, 42, sq, dense_threshold=2000)
    res = O.max_eigenpair_residual(L, r["evals"], r["evecs"])
    assert (r["level"] in (1, 3) and res < 1e-5) or r["level"] in (4, 5)


def test_ring2000_needs_more_than_the_first_pass():
    # test_solver_invariants.rs:337-358: ring C_2000 does not converge in the
    # first 300-iteration pass; a warm restart must fire.  (Which gate fires
    # depends on the LOBPCG iterate, which is not pinned to linfa's.)
    ip, ix, w = ring_graph(2000)
    L, _, sq = _lap(dict(w_indptr=ip, w_indices=ix, w_data=w))
    r = O.lobpcg_solve(L, 2, 42, sq)
    assert r is None or r["passes"] >= 2


# --------------------------------------------------------------- a12
def test_rsvd_ring6_closed_form():
    # rsvd.rs:396-444: ring-6 eigenvalues 1 - cos(2 pi k / 6)
    ip, ix, w = ring_graph(6)
    L, _, sq = _lap(dict(w_indptr=ip, w_indices=ix, w_data=w))
    ev, vec = O.rsvd_solve(L, 2, 42)
    np.testing.assert_allclose(ev, [0.0, 0.5, 0.5], atol=1e-10)
    assert O.max_eigenpair_residual(L, ev, vec) < 1e-10


def test_rsvd_fixed_oversampling_close_to_dense():
    z = load_fixture("blobs_connected_200")
    L, _, _ = _lap(z)
    ev, vec = O.rsvd_solve(L, 2, 42, oversampling=30, nbiter=8)
    np.testing.assert_allclose(ev, z["eig_values"], atol=5e-3)  # test_comp_d_rsvd.rs sanity


# --------------------------------------------------------------- pipeline rows
@pytest.mark.parametrize("name", DATASETS)
def test_components_match_fixture(name):
    z = load_fixture(name)
    labels, nc = O.find_components(z["w_indptr"], z["w_indices"])
    assert nc == int(z["comp_n"])
    # same partition (label names may differ)
    pairs = set(zip(labels.tolist(), z["comp_labels"].tolist()))
    assert len(pairs) == nc


@pytest.mark.parametrize("name", ["blobs_connected_200", "moons_200", "circles_300",
                                  "near_dupes_100", "blobs_connected_2000"])
def test_selection_matches_fixture(name):
    # test_comp_e_selection.rs:31-55 (sign-agnostic 1e-10 on the fixture's own eigvecs)
    z = load_fixture(name)
    sel = O.select_eigenvectors(z["eig_values"], z["eig_vectors"], 2)
    for c in range(2):
        a, b = sel[:, c], z["sel_embedding"][:, c]
        assert min(np.abs(a - b).max(), np.abs(a + b).max()) < 1e-10


@pytest.mark.parametrize("name", DATASETS)
def test_scale_coords_exact(name):
    # scaling.rs:124-135: pre_noise equals scale_coords(embedding, 10) exactly (f32)
    z = load_fixture(name)
    out = O.scale_coords(z["sel_embedding"], 10.0)
    np.testing.assert_array_equal(out, z["scale_pre_noise"])


def test_normalize_signs_argmax_convention():
    c = np.array([[0.1, -3.0], [-2.0, 1.0], [2.0, 0.5]])
    out = O.normalize_signs(c)
    # column 0: first largest |.| is -2.0 -> flipped; column 1: -3.0 -> flipped
    np.testing.assert_array_equal(out, -c)


def test_spectral_init_end_to_end_vs_python_embedding():
    # e2e (test_e2e_validation.rs:23-76): de-scaled embedding close to the
    # Python spectral embedding, sign-agnostic
    z = load_fixture("blobs_connected_200")
    st, out = O.spectral_init(z["w_indptr"], z["w_indices"], z["w_data"], 2, 42, add_noise=False)
    assert st == 0
    ref = z["scale_pre_noise"].astype(np.float64)
    for c in range(2):
        a, b = out[:, c] / 10.0, ref[:, c] / 10.0
        assert min(np.abs(a - b).max(), np.abs(a + b).max()) < 5e-3


def test_spectral_init_disconnected_orthogonal_meta():
    z = load_fixture("disconnected_200")
    st, out = O.spectral_init(z["w_indptr"], z["w_indices"], z["w_data"], 2, 42, add_noise=False)
    assert st == 0 and np.isfinite(out).all()
    assert abs(np.abs(out).max() - 10.0) < 1e-5


def test_spectral_init_errors():
    ip, ix, w = path_graph(2)
    st, _ = O.spectral_init(ip, ix, w, 2, 42)
    assert st == 3  # TooFewNodes
    st, _ = O.spectral_init(np.zeros(1, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float32), 2, 42)
    assert st == 2  # InvalidGraph


# --------------------------------------------------------------- RNG
def test_normal_rng_moments_and_determinism():
    a = O.normal_block(42, 1, 200000)
    b = O.normal_block(42, 1, 200000)
    np.testing.assert_array_equal(a, b)
    assert abs(a.mean()) < 0.01 and abs(a.std() - 1.0) < 0.01
    c = O.normal_block(43, 1, 1000)
    assert not np.allclose(a[:1000], c)


# --------------------------------------------------------------- planted mistakes
def test_planted_transpose_mistake_is_caught():
    # A plausible oracle bug (applying W^T D^-1/2 on the wrong side, i.e. using
    # the transposed degree scaling) changes the Laplacian values and must fail
    # the fixture comparison.
    z = load_fixture("blobs_connected_200")
    deg, sq = O.compute_degrees(z["w_indptr"], z["w_indices"], z["w_data"], 0)
    isd = 1.0 / sq
    wrong = O.build_laplacian(z["w_indptr"], z["w_indices"], z["w_data"], isd ** 2)
    assert np.abs(wrong[2] - z["l_data"]).max() > 1e-3


def test_planted_sign_mistake_in_chebyshev_is_caught():
    lam = np.linspace(0.0, 2.0, 11)
    r = np.ones((11, 1))
    good = O.chebyshev_filter(_diag_csr(lam), r, 0.1, 2.0, 3)
    bad = O.chebyshev_filter(_diag_csr(-lam), r, 0.1, 2.0, 3)  # wrong-sign operator
    t3 = npcheb.chebval((lam - 1.05) / 0.95, [0, 0, 0, 1])
    np.testing.assert_allclose(good[:, 0], t3, atol=1e-12)
    assert np.abs(bad[:, 0] - t3).max() > 1.0
