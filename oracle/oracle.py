"""ctypes binding of the FP64 CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the benchmark's
cpu_baseline leg load this as the checker.  The product library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")


class LobpcgStats(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("spmm_calls", C.c_int64),
                ("restarts", C.c_int), ("passes", C.c_int), ("converged", C.c_int)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_set_threads.argtypes = [C.c_int]
        L.or_philox4x32_10.argtypes = [_u32p, C.c_uint32, C.c_uint32, _u32p]
        L.or_normal.restype = C.c_double
        L.or_normal.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64]
        L.or_fill_normal.argtypes = [C.c_uint64, C.c_uint32, C.c_size_t, _f64p]
        L.or_normal_pairs_f32.argtypes = [C.c_uint64, C.c_uint32, C.c_size_t, _f32p]
        L.or_compute_degrees.argtypes = [C.c_int64, _i64p, _i32p, _f32p, C.c_int, _f64p, _f64p]
        L.or_laplacian_nnz.restype = C.c_int64
        L.or_laplacian_nnz.argtypes = [C.c_int64, _i64p, _i32p]
        L.or_build_laplacian.argtypes = [C.c_int64, _i64p, _i32p, _f32p, _f64p, _i64p, _i32p, _f64p]
        L.or_spmm.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_int, _f64p]
        L.or_eigenpair_residual.restype = C.c_double
        L.or_eigenpair_residual.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_double]
        L.or_max_eigenpair_residual.restype = C.c_double
        L.or_max_eigenpair_residual.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, _f64p, C.c_int]
        L.or_sym_eig.argtypes = [C.c_int, _f64p, _f64p, _f64p]
        L.or_gen_sym_eig.restype = C.c_int
        L.or_gen_sym_eig.argtypes = [C.c_int, _f64p, _f64p, _f64p, _f64p]
        L.or_cholesky.restype = C.c_int
        L.or_cholesky.argtypes = [C.c_int, _f64p]
        L.or_dense_evd.restype = C.c_int
        L.or_dense_evd.argtypes = [C.c_int64, _i64p, _i32p, _f64p, C.c_int, _f64p, _f64p]
        L.or_chebyshev_filter.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_int,
                                          C.c_double, C.c_double, C.c_int, _f64p]
        L.or_rayleigh_ritz.restype = C.c_int
        L.or_rayleigh_ritz.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_int, _f64p]
        L.or_per_vector_residuals.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, _f64p,
                                              C.c_int, _f64p]
        L.or_lobpcg_core.restype = C.c_int
        L.or_lobpcg_core.argtypes = [C.c_int64, _i64p, _i32p, _f64p, C.c_double, _f64p, C.c_int,
                                     C.c_double, C.c_int, _f64p, _f64p,
                                     C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.or_lobpcg_solve.restype = C.c_int
        L.or_lobpcg_solve.argtypes = [C.c_int64, _i64p, _i32p, _f64p, C.c_int, C.c_uint64, C.c_int,
                                      _f64p, C.c_double, C.c_void_p, _f64p, _f64p,
                                      C.POINTER(LobpcgStats)]
        L.or_rsvd_solve.argtypes = [C.c_int64, _i64p, _i32p, _f64p, C.c_int, C.c_uint64, C.c_int,
                                    C.c_int, _f64p, _f64p]
        L.or_solve_eigenproblem.restype = C.c_int
        L.or_solve_eigenproblem.argtypes = [C.c_int64, _i64p, _i32p, _f64p, C.c_int, C.c_uint64,
                                            _f64p, C.c_int64, C.c_double, _f64p, _f64p,
                                            C.POINTER(LobpcgStats)]
        L.or_find_components.restype = C.c_int64
        L.or_find_components.argtypes = [C.c_int64, _i64p, _i32p, _i64p]
        L.or_select_eigenvectors.argtypes = [_f64p, _f64p, C.c_int64, C.c_int, C.c_int, _f64p]
        L.or_normalize_signs.argtypes = [_f64p, C.c_int64, C.c_int]
        L.or_scale_coords.restype = C.c_int
        L.or_scale_coords.argtypes = [_f64p, C.c_int64, C.c_int, C.c_double, _f32p]
        L.or_spectral_init.restype = C.c_int
        L.or_spectral_init.argtypes = [C.c_int64, _i64p, _i32p, _f32p, C.c_int, C.c_uint64,
                                       C.c_void_p, C.c_int64, C.c_int, C.c_int64, C.c_double,
                                       C.c_int, _f32p]
        if hasattr(L, "or_mpnn_forward"):
            _bind_mpnn(L)
        L.or_knn_exact.argtypes = [C.c_int64, C.c_int, _f64p, C.c_int, _i32p, _f32p]
        L.or_smooth_knn_dist.argtypes = [C.c_int64, C.c_int, _f32p, _f32p, _f32p]
        L.or_membership.argtypes = [C.c_int64, C.c_int, _i32p, _f32p, _f32p, _f32p, _f32p]
        L.or_fuzzy_union.restype = C.c_int64
        L.or_fuzzy_union.argtypes = [C.c_int64, C.c_int, _i32p, _f32p, C.c_int, C.c_void_p,
                                     C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _bind_mpnn(L):
    L.or_mpnn_param_count.restype = C.c_int64
    L.or_mpnn_param_count.argtypes = []
    L.or_mpnn_features.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_uint64, _f64p]
    L.or_mpnn_forward.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                  _f64p]
    L.or_mpnn_loss.restype = C.c_double
    L.or_mpnn_loss.argtypes = [C.c_int64, _f64p, _f64p, C.c_void_p]
    L.or_mpnn_backward.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                   _f64p, _f64p]


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


# ------------------------------------------------------------------ helpers
def _csr_w(indptr, indices, data):
    return (np.ascontiguousarray(indptr, dtype=np.int64),
            np.ascontiguousarray(indices, dtype=np.int32),
            np.ascontiguousarray(data, dtype=np.float32))


def _csr_l(indptr, indices, data):
    return (np.ascontiguousarray(indptr, dtype=np.int64),
            np.ascontiguousarray(indices, dtype=np.int32),
            np.ascontiguousarray(data, dtype=np.float64))


def philox4x32_10(ctr, key0: int, key1: int) -> np.ndarray:
    """The raw Philox4x32-10 output block for one 4-word counter."""
    c = np.ascontiguousarray(ctr, np.uint32)
    out = np.empty(4, np.uint32)
    lib().or_philox4x32_10(c, key0, key1, out)
    return out


def normal_at(seed: int, stream: int, index: int) -> float:
    return lib().or_normal(seed, stream, index)


def normal_block(seed: int, stream: int, count: int) -> np.ndarray:
    out = np.empty(count, np.float64)
    lib().or_fill_normal(seed, stream, count, out)
    return out


def normal_pairs_f32(seed: int, stream: int, npairs: int) -> np.ndarray:
    """Pairs 0 .. npairs-1 of the FP32 start-block stream, float32 [npairs, 2]
    (cosine and sine output of each pair's Box-Muller transform)."""
    out = np.empty(2 * npairs, np.float32)
    lib().or_normal_pairs_f32(seed, stream, npairs, out)
    return out.reshape(npairs, 2)


def compute_degrees(indptr, indices, data, mode: int = 0):
    ip, ix, w = _csr_w(indptr, indices, data)
    n = ip.size - 1
    deg = np.empty(n, np.float64)
    sq = np.empty(n, np.float64)
    lib().or_compute_degrees(n, ip, ix, w, int(mode), deg, sq)
    return deg, sq


def build_laplacian(indptr, indices, data, inv_sqrt_deg):
    ip, ix, w = _csr_w(indptr, indices, data)
    n = ip.size - 1
    nnz = lib().or_laplacian_nnz(n, ip, ix)
    lip = np.empty(n + 1, np.int64)
    lix = np.empty(nnz, np.int32)
    ld = np.empty(nnz, np.float64)
    lib().or_build_laplacian(n, ip, ix, w, np.ascontiguousarray(inv_sqrt_deg, np.float64), lip, lix, ld)
    m = int(lip[-1])
    return lip, lix[:m].copy(), ld[:m].copy()


def laplacian_from_graph(indptr, indices, data, mode: int = 0):
    deg, sq = compute_degrees(indptr, indices, data, mode)
    isd = np.where(sq > 0, 1.0 / np.where(sq > 0, sq, 1.0), 0.0)
    return build_laplacian(indptr, indices, data, isd), deg, sq


def spmm(L, x):
    ip, ix, d = _csr_l(*L)
    x = np.ascontiguousarray(x, np.float64)
    squeeze = x.ndim == 1
    if squeeze:
        x = x[:, None]
    y = np.empty_like(x)
    lib().or_spmm(ip.size - 1, ip, ix, d, x, x.shape[1], y)
    return y[:, 0] if squeeze else y


def max_eigenpair_residual(L, evals, evecs):
    ip, ix, d = _csr_l(*L)
    evecs = np.ascontiguousarray(evecs, np.float64)
    return lib().or_max_eigenpair_residual(ip.size - 1, ip, ix, d, evecs,
                                           np.ascontiguousarray(evals, np.float64), evecs.shape[1])


def sym_eig(a):
    a = np.ascontiguousarray(a, np.float64)
    n = a.shape[0]
    w = np.empty(n)
    v = np.empty((n, n))
    lib().or_sym_eig(n, a, w, v)
    return w, v


def gen_sym_eig(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    n = a.shape[0]
    w = np.empty(n)
    c = np.empty((n, n))
    st = lib().or_gen_sym_eig(n, a, b, w, c)
    if st != 0:
        raise np.linalg.LinAlgError("B not positive definite")
    return w, c


def dense_evd(L, k):
    ip, ix, d = _csr_l(*L)
    n = ip.size - 1
    ev = np.empty(k)
    vec = np.empty((n, k))
    st = lib().or_dense_evd(n, ip, ix, d, k, ev, vec)
    if st != 0:
        raise RuntimeError("dense_evd failed")
    return ev, vec


def chebyshev_filter(L, r, a=0.1, b=2.0, degree=8):
    ip, ix, d = _csr_l(*L)
    r = np.ascontiguousarray(r, np.float64)
    out = np.empty_like(r)
    lib().or_chebyshev_filter(ip.size - 1, ip, ix, d, r, r.shape[1], a, b, degree, out)
    return out


def rayleigh_ritz(L, x):
    ip, ix, d = _csr_l(*L)
    x = np.array(x, np.float64, order="C", copy=True)
    ev = np.empty(x.shape[1])
    st = lib().or_rayleigh_ritz(ip.size - 1, ip, ix, d, x, x.shape[1], ev)
    if st != 0:
        raise RuntimeError("rayleigh_ritz failed")
    return ev, x


def per_vector_residuals(L, evals, x):
    ip, ix, d = _csr_l(*L)
    x = np.ascontiguousarray(x, np.float64)
    res = np.empty(x.shape[1])
    lib().or_per_vector_residuals(ip.size - 1, ip, ix, d, np.ascontiguousarray(evals, np.float64),
                                  x, x.shape[1], res)
    return res


def lobpcg_core(L, x0, tol=1e-5, maxiter=300, shift=0.0):
    ip, ix, d = _csr_l(*L)
    x = np.array(x0, np.float64, order="C", copy=True)
    m = x.shape[1]
    ev = np.empty(m)
    rn = np.empty(m)
    it = C.c_int64(0)
    calls = C.c_int64(0)
    conv = lib().or_lobpcg_core(ip.size - 1, ip, ix, d, shift, x, m, tol, maxiter, ev, rn,
                                C.byref(it), C.byref(calls))
    return dict(converged=bool(conv), evals=ev, evecs=x, rnorm=rn, iterations=it.value,
                spmm_calls=calls.value)


def lobpcg_solve(L, n_components, seed, sqrt_deg, regularize=False, tol=1e-5, x0=None):
    ip, ix, d = _csr_l(*L)
    n = ip.size - 1
    k = n_components + 1
    ev = np.empty(k)
    vec = np.empty((n, k))
    st = LobpcgStats()
    x0p = None
    if x0 is not None:
        x0 = np.ascontiguousarray(x0, np.float64)
        x0p = x0.ctypes.data
    ok = lib().or_lobpcg_solve(n, ip, ix, d, n_components, seed, int(regularize),
                               np.ascontiguousarray(sqrt_deg, np.float64), tol, x0p, ev, vec,
                               C.byref(st))
    if not ok:
        return None
    return dict(evals=ev, evecs=vec, iterations=st.iterations, spmm_calls=st.spmm_calls,
                restarts=st.restarts, passes=st.passes, converged=bool(st.converged))


def rsvd_solve(L, n_components, seed, oversampling=-1, nbiter=2):
    ip, ix, d = _csr_l(*L)
    n = ip.size - 1
    k = n_components + 1
    ev = np.empty(k)
    vec = np.empty((n, k))
    lib().or_rsvd_solve(n, ip, ix, d, n_components, seed, oversampling, nbiter, ev, vec)
    return ev, vec


def solve_eigenproblem(L, n_components, seed, sqrt_deg, dense_threshold=2000, tol=1e-5):
    ip, ix, d = _csr_l(*L)
    n = ip.size - 1
    k = n_components + 1
    ev = np.empty(k)
    vec = np.empty((n, k))
    st = LobpcgStats()
    level = lib().or_solve_eigenproblem(n, ip, ix, d, n_components, seed,
                                        np.ascontiguousarray(sqrt_deg, np.float64),
                                        dense_threshold, tol, ev, vec, C.byref(st))
    return dict(evals=ev, evecs=vec, level=level, iterations=st.iterations,
                restarts=st.restarts)


def find_components(indptr, indices):
    ip = np.ascontiguousarray(indptr, np.int64)
    ix = np.ascontiguousarray(indices, np.int32)
    labels = np.empty(ip.size - 1, np.int64)
    nc = lib().or_find_components(ip.size - 1, ip, ix, labels)
    return labels, int(nc)


def select_eigenvectors(evals, evecs, n_components):
    evecs = np.ascontiguousarray(evecs, np.float64)
    out = np.empty((evecs.shape[0], n_components))
    lib().or_select_eigenvectors(np.ascontiguousarray(evals, np.float64), evecs, evecs.shape[0],
                                 evecs.shape[1], n_components, out)
    return out


def normalize_signs(coords):
    c = np.array(coords, np.float64, order="C", copy=True)
    lib().or_normalize_signs(c, c.shape[0], c.shape[1])
    return c


def scale_coords(coords, max_coord=10.0):
    c = np.ascontiguousarray(coords, np.float64)
    out = np.empty(c.shape, np.float32)
    if lib().or_scale_coords(c, c.shape[0], c.shape[1], max_coord, out) != 0:
        raise ValueError("all-zero embedding")
    return out


def spectral_init(indptr, indices, data, n_components, seed, raw_data=None, mode=0,
                  dense_threshold=2000, tol=1e-5, add_noise=True):
    ip, ix, w = _csr_w(indptr, indices, data)
    n = ip.size - 1
    out = np.empty((n, n_components), np.float32)
    rawp, nf = None, 0
    if raw_data is not None:
        raw_data = np.ascontiguousarray(raw_data, np.float32)
        rawp, nf = raw_data.ctypes.data, raw_data.shape[1]
    st = lib().or_spectral_init(n, ip, ix, w, n_components, seed, rawp, nf, mode,
                                dense_threshold, tol, int(add_noise), out)
    return st, out


# ---------------------------------------------------------------- warm-start network
MPNN_HID, MPNN_NF, MPNN_NOUT = 64, 16, 8


def adjacency_from_graph(indptr, indices, data, mode: int = 0):
    """Off-diagonal normalised adjacency A_hat = I - L as CSR (ip, ix, a) and degrees."""
    (lip, lix, ld), deg, _ = laplacian_from_graph(indptr, indices, data, mode)
    keep = lix != np.repeat(np.arange(lip.size - 1), np.diff(lip))
    a = -ld[keep]
    counts = np.add.reduceat(keep.astype(np.int64), lip[:-1]) if lix.size else np.zeros(lip.size - 1, np.int64)
    counts[np.diff(lip) == 0] = 0
    ip = np.zeros(lip.size, np.int64)
    ip[1:] = np.cumsum(counts)
    return (ip, np.ascontiguousarray(lix[keep], np.int32), np.ascontiguousarray(a)), deg


def mpnn_param_count() -> int:
    return int(lib().or_mpnn_param_count())


def mpnn_features(adj, deg, seed: int) -> np.ndarray:
    ip, ix, a = adj
    n = ip.size - 1
    f = np.empty((n, MPNN_NF))
    lib().or_mpnn_features(n, ip, ix, a, np.ascontiguousarray(deg, np.float64), seed, f)
    return f


def mpnn_forward(adj, f, params):
    ip, ix, a = adj
    n = ip.size - 1
    h = np.zeros((4, n, MPNN_HID))
    av = np.zeros((3, n, MPNN_HID))
    x = np.empty((n, MPNN_NOUT))
    lib().or_mpnn_forward(n, ip, ix, a, np.ascontiguousarray(f, np.float64),
                          np.ascontiguousarray(params, np.float64), h, av, x)
    return x, h, av


def mpnn_loss(x, v, grad: bool = True):
    x = np.ascontiguousarray(x, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    dx = np.empty_like(x) if grad else None
    loss = lib().or_mpnn_loss(x.shape[0], x, v, dx.ctypes.data if grad else None)
    return (loss, dx) if grad else loss


def mpnn_backward(adj, f, params, h, av, dx):
    ip, ix, a = adj
    n = ip.size - 1
    g = np.empty(mpnn_param_count())
    lib().or_mpnn_backward(n, ip, ix, a, np.ascontiguousarray(f, np.float64),
                           np.ascontiguousarray(params, np.float64), h, av,
                           np.ascontiguousarray(dx, np.float64), g)
    return g


# ------------------------------------------------------------------ fuzzy k-NN graph
def knn_exact(x, k: int):
    """Exact k-NN, self excluded: (indices int32 [n, k], distances float32 [n, k])."""
    x = np.ascontiguousarray(x, np.float64)
    n, d = x.shape
    idx = np.empty((n, k), np.int32)
    dist = np.empty((n, k), np.float32)
    lib().or_knn_exact(n, d, x, k, idx, dist)
    return idx, dist


def smooth_knn_dist(dist):
    """umap smooth_knn_dist: (sigmas, rhos), float32 [n]."""
    dist = np.ascontiguousarray(dist, np.float32)
    n, k = dist.shape
    sig = np.empty(n, np.float32)
    rho = np.empty(n, np.float32)
    lib().or_smooth_knn_dist(n, k, dist, sig, rho)
    return sig, rho


def membership(idx, dist, sigmas, rhos):
    """umap compute_membership_strengths: float32 [n, k] values of A."""
    idx = np.ascontiguousarray(idx, np.int32)
    dist = np.ascontiguousarray(dist, np.float32)
    n, k = idx.shape
    vals = np.empty((n, k), np.float32)
    lib().or_membership(n, k, idx, dist, np.ascontiguousarray(sigmas, np.float32),
                        np.ascontiguousarray(rhos, np.float32), vals)
    return vals


def fuzzy_union(idx, vals, n_epochs: int = 0):
    """W = A + A^T - A o A^T (zeros dropped), pruned below max/n_epochs when
    n_epochs > 0: CSR (int64 indptr, int32 indices, float32 data)."""
    idx = np.ascontiguousarray(idx, np.int32)
    vals = np.ascontiguousarray(vals, np.float32)
    n, k = idx.shape
    nnz = lib().or_fuzzy_union(n, k, idx, vals, int(n_epochs), None, None, None)
    ip = np.empty(n + 1, np.int64)
    ix = np.empty(nnz, np.int32)
    w = np.empty(nnz, np.float32)
    lib().or_fuzzy_union(n, k, idx, vals, int(n_epochs), ip.ctypes.data, ix.ctypes.data,
                         w.ctypes.data)
    return ip, ix, w


def fuzzy_graph(x, k: int = 15, n_epochs: int = -1):
    """Steps 1-5a from raw points; n_epochs -1 = the fixtures' rule (500 for
    n <= 10000, else 200), 0 = no pruning."""
    n = np.asarray(x).shape[0]
    if n_epochs < 0:
        n_epochs = 500 if n <= 10000 else 200
    idx, dist = knn_exact(x, k)
    sig, rho = smooth_knn_dist(dist)
    return fuzzy_union(idx, membership(idx, dist, sig, rho), n_epochs)
