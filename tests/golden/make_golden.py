"""Extract the golden vectors this project's tests need from the reference's fixture data.

Only *data files* the reference's own tests hold are read (``.npz`` / ``.npy`` /
``.json``, loaded with ``allow_pickle=False``); no reference code is imported or run.
The output is one compressed ``.npz`` per dataset under ``tests/golden/fixtures/``
plus the Rust-produced eigenvector blocks under ``tests/golden/rust_vectors/``, so
that the tests (and the GPU runs, where the reference tree does not exist) read
committed data only.

Sources (relative to the reference checkout):
  tests/fixtures/<dataset>/step5a_pruned.npz     input graph W (CSR f32, int32 idx)
  tests/fixtures/<dataset>/comp_a_degrees.npz    degrees, sqrt_deg (f64)
  tests/fixtures/<dataset>/comp_b_laplacian.npz  L = I - D^-1/2 W D^-1/2 (CSR f64)
  tests/fixtures/<dataset>/comp_c_components.npz component labels
  tests/fixtures/<dataset>/comp_d_eigensolver.npz scipy eigsh eigenpairs (k=3)
  tests/fixtures/<dataset>/comp_e_selection.npz  selected columns
  tests/fixtures/<dataset>/comp_f_scaling.npz    pre_noise scaled coordinates
  tests/fixtures/<dataset>/step0_raw_data.npz .. step5a_pruned_exact.npz
                                                 the fuzzy k-NN graph steps (exact k-NN path)
                                                 -> tests/golden/graph/<dataset>.npz
  research/2026-03-27-computemode-simd-spmv-parity/results/eigenvectors/*.npy
  research/2026-03-25-ci-test-optimization/results/rq3_threshold/*_n200_eigvecs.npy

Run (in the development container only):  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

REF = os.environ.get("SPECTRAL_REFERENCE_ROOT", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
OUT_FIX = os.path.join(HERE, "fixtures")
OUT_RUST = os.path.join(HERE, "rust_vectors")

DATASETS = [
    "blobs_50", "blobs_500", "blobs_5000", "moons_200", "circles_300",
    "near_dupes_100", "disconnected_200", "blobs_connected_200", "blobs_connected_2000",
]


def _load(path):
    return np.load(path, allow_pickle=False)


def extract_dataset(name: str) -> dict:
    d = os.path.join(REF, "tests", "fixtures", name)
    out = {}
    w = _load(os.path.join(d, "step5a_pruned.npz"))
    out["w_indptr"] = w["indptr"].astype(np.int64)
    out["w_indices"] = w["indices"].astype(np.int32)
    out["w_data"] = w["data"].astype(np.float32)
    out["w_shape"] = w["shape"].astype(np.int64)
    a = _load(os.path.join(d, "comp_a_degrees.npz"))
    out["degrees"] = a["degrees"]
    out["sqrt_deg"] = a["sqrt_deg"]
    b = _load(os.path.join(d, "comp_b_laplacian.npz"))
    out["l_indptr"] = b["indptr"].astype(np.int64)
    out["l_indices"] = b["indices"].astype(np.int32)
    out["l_data"] = b["data"]
    c = _load(os.path.join(d, "comp_c_components.npz"))
    out["comp_labels"] = c["labels"]
    out["comp_n"] = np.asarray(c["n_components"])
    e = _load(os.path.join(d, "comp_d_eigensolver.npz"))
    out["eig_values"] = e["eigenvalues"]
    out["eig_vectors"] = e["eigenvectors"]
    out["eig_residuals"] = e["residuals"]
    s = _load(os.path.join(d, "comp_e_selection.npz"))
    out["sel_order"] = s["order"]
    out["sel_embedding"] = s["embedding"]
    f = _load(os.path.join(d, "comp_f_scaling.npz"))
    out["scale_pre_noise"] = f["pre_noise"]
    x = _load(os.path.join(d, "step0_raw_data.npz"))
    out["raw_X"] = x["X"].astype(np.float32)
    with open(os.path.join(d, "meta.json")) as fh:
        out["meta_json"] = np.asarray(json.dumps(json.load(fh)))
    return out


def _csr(z):
    """CSR arrays of a scipy save_npz file (format 'csr'), read as plain data."""
    assert z["format"].item() in (b"csr", "csr")
    return (z["indptr"].astype(np.int64), z["indices"].astype(np.int32),
            z["data"].astype(np.float32))


def extract_graph_steps(name: str) -> dict:
    """Steps 0-5a of the exact k-NN path (generate_fixtures.py:451-471)."""
    d = os.path.join(REF, "tests", "fixtures", name)
    out = {}
    x = _load(os.path.join(d, "step0_raw_data.npz"))
    out["x"] = x["X"].astype(np.float64)
    s1 = _load(os.path.join(d, "step1_knn_exact.npz"))
    out["knn_indices"] = s1["knn_indices"].astype(np.int32)
    out["knn_dists"] = s1["knn_dists"].astype(np.float32)
    out["n_neighbors"] = np.asarray(s1["n_neighbors"], np.int32)
    s2 = _load(os.path.join(d, "step2_smooth_knn_exact.npz"))
    out["sigmas"] = s2["sigmas"].astype(np.float32)
    out["rhos"] = s2["rhos"].astype(np.float32)
    s3 = _load(os.path.join(d, "step3_membership_exact.npz"))
    assert s3["format"].item() in (b"coo", "coo")
    out["s3_row"] = s3["row"].astype(np.int32)
    out["s3_col"] = s3["col"].astype(np.int32)
    out["s3_data"] = s3["data"].astype(np.float32)
    for tag, fname in (("s4", "step4_symmetrized_exact.npz"), ("s5", "step5a_pruned_exact.npz")):
        ip, ix, dv = _csr(_load(os.path.join(d, fname)))
        out[f"{tag}_indptr"], out[f"{tag}_indices"], out[f"{tag}_data"] = ip, ix, dv
    return out


def main() -> int:
    os.makedirs(os.path.join(HERE, "graph"), exist_ok=True)
    for name in DATASETS:
        arrays = extract_graph_steps(name)
        np.savez_compressed(os.path.join(HERE, "graph", f"{name}.npz"), **arrays)
        print(f"{name}: graph steps, n={arrays['x'].shape[0]} d={arrays['x'].shape[1]} "
              f"nnz(step4)={arrays['s4_data'].size} nnz(step5a)={arrays['s5_data'].size}")
    os.makedirs(OUT_FIX, exist_ok=True)
    os.makedirs(OUT_RUST, exist_ok=True)
    for name in DATASETS:
        arrays = extract_dataset(name)
        np.savez_compressed(os.path.join(OUT_FIX, f"{name}.npz"), **arrays)
        print(f"{name}: n={arrays['w_shape'][0]} nnz(W)={arrays['w_data'].size}")
    rust = {
        "blobs_connected_2000_scalar.npy": "research/2026-03-27-computemode-simd-spmv-parity/results/eigenvectors/blobs_connected_2000_scalar.npy",
        "blobs_5000_scalar.npy": "research/2026-03-27-computemode-simd-spmv-parity/results/eigenvectors/blobs_5000_scalar.npy",
        "ring200_lobpcg_eigvecs.npy": "research/2026-03-25-ci-test-optimization/results/rq3_threshold/lobpcg_n200_eigvecs.npy",
        "ring200_dense_eigvecs.npy": "research/2026-03-25-ci-test-optimization/results/rq3_threshold/dense_n200_eigvecs.npy",
    }
    for dst, src in rust.items():
        arr = _load(os.path.join(REF, src))
        np.save(os.path.join(OUT_RUST, dst), arr, allow_pickle=False)
        print(f"{dst}: {arr.shape} {arr.dtype}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
