"""Python mirror of spectral-init's public interface over the MI355X C ABI.

Names follow the reference crate: ``spectral_init``, ``SpectralInitConfig``,
``ComputeMode``, ``SpectralError`` (``ConvergenceFailure``, ``InvalidGraph``,
``TooFewNodes``), plus the seams the reference exposes for testing
(``compute_degrees``/``build_normalized_laplacian`` -> ``Laplacian``,
``CsrOperator.apply`` -> ``Laplacian.spmm``, ``lobpcg_solve``,
``solve_eigenproblem``, ``rsvd_solve``, ``dense_evd``,
``max_eigenpair_residual``).

The library is ``lib/libspectral_mi355x.so`` (built by ``build.py``).  Loading
it does not need a GPU; every compute call does, and fails loudly without one.
Device blocks are passed as torch CUDA tensors (torch is only the allocator
here) or raw integer device addresses.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libspectral_mi355x.so")

SI_OK, SI_ERR_CONVERGENCE, SI_ERR_INVALID_GRAPH, SI_ERR_TOO_FEW_NODES = 0, 1, 2, 3
SI_ERR_HIP, SI_ERR_INTERNAL, SI_ERR_INVALID_ARGUMENT, SI_ERR_UNSUPPORTED = 4, 5, 6, 7
SI_F32, SI_F64 = 0, 1
SI_BUILD_LOCALITY_ORDER, SI_BUILD_RCM_ORDER = 1, 2

# every symbol the public and testing headers declare
EXPORTED = [
    "si_version", "si_config_default", "si_context_create", "si_context_destroy",
    "si_last_error", "si_laplacian_build", "si_laplacian_build_ex", "si_laplacian_destroy",
    "si_laplacian_info",
    "si_laplacian_export", "si_spmm", "si_chebyshev_filter", "si_max_eigenpair_residual",
    "si_max_eigenpair_residual_noalloc", "si_operator_apply_host", "si_spmv_csr",
    "si_lobpcg_solve", "si_lobpcg_sinv_solve", "si_solve_eigenproblem", "si_rsvd_solve",
    "si_dense_evd",
    "si_spectral_init",
    "si_test_gram", "si_test_block_mul", "si_test_small_sym_eig", "si_test_small_gen_eig",
    "si_test_fill_normal", "si_test_fill_normal_f32", "si_test_staged_roundtrip", "si_test_fast_solve", "si_test_batch_graphs", "si_test_batch_plan",
    "si_test_sinv_dense",
    "si_test_spmm_shapes",
    "si_test_band_solve", "si_test_rcm_order", "si_test_knn_path",
    "si_test_laplacian_storage", "si_test_components_device", "si_test_rcm_order_device",
    "si_test_output_stage",
    "si_profile_spmm_begin", "si_profile_spmm_end", "si_profile_gather_probe",
    "si_mpnn_param_count", "si_mpnn_create", "si_mpnn_destroy", "si_mpnn_get_params",
    "si_mpnn_set_params", "si_mpnn_features", "si_mpnn_forward", "si_mpnn_gradient",
    "si_mpnn_train_step", "si_mpnn_warm_start",
    "si_knn_exact", "si_smooth_knn_dist", "si_membership_strengths", "si_fuzzy_union",
    "si_fuzzy_graph_build", "si_graph_info", "si_graph_copy", "si_graph_destroy",
]


class _SpmmProfile(C.Structure):
    _fields_ = [("launches", C.c_int64), ("bytes", C.c_double), ("ms", C.c_double),
                ("launches_f32", C.c_int64), ("bytes_f32", C.c_double), ("ms_f32", C.c_double)]


class ComputeMode(enum.IntEnum):
    """src/config.rs:3-14"""
    PythonCompat = 0
    RustNative = 1


class Solver(enum.IntEnum):
    Fast = 0
    PlainLobpcg = 1


class _Config(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("compute_mode", C.c_int32), ("solver", C.c_int32),
                ("block_size", C.c_int32), ("cheb_degree", C.c_int32), ("max_outer", C.c_int32),
                ("tol", C.c_double), ("dense_threshold", C.c_int64), ("warm_start", C.c_void_p),
                ("batch_components", C.c_int32)]


class _Report(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("level", C.c_int32), ("restarts", C.c_int32),
                ("block_size", C.c_int32), ("iterations", C.c_int64), ("spmm_calls", C.c_int64),
                ("max_residual", C.c_double), ("eigenvalues", C.c_double * 8),
                ("solve_ms", C.c_double), ("laplacian_ms", C.c_double),
                ("passes", C.c_int32), ("reserved", C.c_int32), ("level_ms", C.c_double * 6)]


@dataclass
class SpectralInitConfig:
    """SpectralInitConfig (src/config.rs:19-25) plus the device knobs."""
    compute_mode: ComputeMode = ComputeMode.PythonCompat
    solver: Solver = Solver.Fast
    block_size: int = 0
    cheb_degree: int = 0
    max_outer: int = 60
    tol: float = 1e-5
    dense_threshold: int = 2000
    warm_start: Optional[int] = None  # device address of an [n x (n_components+1)] f64 block
    batch_components: bool = True     # disconnected graphs: one batched device solve

    def _c(self) -> _Config:
        c = _Config()
        c.struct_size = C.sizeof(_Config)
        c.compute_mode = int(self.compute_mode)
        c.solver = int(self.solver)
        c.block_size = int(self.block_size)
        c.cheb_degree = int(self.cheb_degree)
        c.max_outer = int(self.max_outer)
        c.tol = float(self.tol)
        c.dense_threshold = int(self.dense_threshold)
        c.warm_start = self.warm_start
        c.batch_components = 1 if self.batch_components else 0
        return c


@dataclass
class Report:
    level: int = -1
    restarts: int = 0
    block_size: int = 0
    iterations: int = 0
    spmm_calls: int = 0
    max_residual: float = float("nan")
    eigenvalues: list = field(default_factory=list)
    solve_ms: float = 0.0
    laplacian_ms: float = 0.0
    passes: int = 0
    level_ms: list = field(default_factory=list)

    @staticmethod
    def _from(r: _Report, k: int = 8) -> "Report":
        return Report(r.level, r.restarts, r.block_size, r.iterations, r.spmm_calls,
                      r.max_residual, list(r.eigenvalues)[:k], r.solve_ms, r.laplacian_ms,
                      r.passes, list(r.level_ms))


class SpectralError(Exception):
    """SpectralError (src/lib.rs:167-182)."""
    code = SI_ERR_INTERNAL


class ConvergenceFailure(SpectralError):
    code = SI_ERR_CONVERGENCE


class InvalidGraph(SpectralError):
    code = SI_ERR_INVALID_GRAPH


class TooFewNodes(SpectralError):
    code = SI_ERR_TOO_FEW_NODES


class DeviceError(SpectralError):
    code = SI_ERR_HIP


class Unsupported(SpectralError):
    code = SI_ERR_UNSUPPORTED


_ERRORS = {SI_ERR_CONVERGENCE: ConvergenceFailure, SI_ERR_INVALID_GRAPH: InvalidGraph,
           SI_ERR_TOO_FEW_NODES: TooFewNodes, SI_ERR_HIP: DeviceError,
           SI_ERR_INVALID_ARGUMENT: ValueError, SI_ERR_UNSUPPORTED: Unsupported}

_lib = None


def library_path() -> str:
    return LIB_PATH


def lib():
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run spectral-init_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, u64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double
        L.si_version.restype = C.c_char_p
        L.si_last_error.restype = C.c_char_p
        L.si_last_error.argtypes = [vp]
        L.si_config_default.argtypes = [C.POINTER(_Config)]
        L.si_context_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.si_context_destroy.argtypes = [vp]
        L.si_laplacian_build.argtypes = [vp, i64, vp, vp, vp, C.c_int, i32, C.POINTER(vp)]
        L.si_laplacian_build_ex.argtypes = [vp, i64, vp, vp, vp, C.c_int, i32, C.c_uint32,
                                            C.POINTER(vp)]
        L.si_laplacian_destroy.argtypes = [vp]
        L.si_laplacian_info.argtypes = [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64),
                                        C.POINTER(i32)]
        L.si_laplacian_export.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.POINTER(i64)]
        L.si_spmm.argtypes = [vp, vp, i32, vp, vp, i32, i32]
        L.si_chebyshev_filter.argtypes = [vp, vp, vp, i32, dbl, dbl, i32, vp]
        L.si_max_eigenpair_residual.argtypes = [vp, vp, vp, vp, i32, C.POINTER(dbl)]
        L.si_max_eigenpair_residual_noalloc.argtypes = [vp, vp, vp, vp, i32, C.POINTER(dbl)]
        L.si_operator_apply_host.argtypes = [vp, vp, vp, vp, i32]
        L.si_spmv_csr.argtypes = [vp, i64, vp, vp, vp, i64, vp, i64, vp]
        L.si_lobpcg_solve.argtypes = [vp, vp, i32, u64, i32, dbl, vp, vp, C.POINTER(_Report)]
        L.si_lobpcg_sinv_solve.argtypes = [vp, vp, i32, u64, vp, vp, C.POINTER(_Report)]
        L.si_solve_eigenproblem.argtypes = [vp, vp, i32, u64, C.POINTER(_Config), vp, vp,
                                            C.POINTER(_Report)]
        L.si_rsvd_solve.argtypes = [vp, vp, i32, u64, i32, i32, vp, vp, C.POINTER(_Report)]
        L.si_dense_evd.argtypes = [vp, vp, i32, vp, vp]
        L.si_spectral_init.argtypes = [vp, i64, vp, vp, vp, i32, u64, vp, i64, C.POINTER(_Config),
                                       vp, C.POINTER(_Report)]
        L.si_test_gram.argtypes = [vp, i32, i64, vp, i32, i32, vp, i32, i32, vp]
        L.si_test_block_mul.argtypes = [vp, i32, i64, vp, i32, i32, vp, i32, vp, i32]
        L.si_test_small_sym_eig.argtypes = [vp, i32, vp, vp, vp]
        L.si_test_small_gen_eig.argtypes = [vp, i32, vp, vp, vp, vp]
        L.si_test_fill_normal.argtypes = [vp, vp, i64, i32, u64, C.c_uint32]
        L.si_test_fill_normal_f32.argtypes = [vp, vp, i64, i32, u64, C.c_uint32]
        L.si_test_staged_roundtrip.argtypes = [vp, vp, vp, i64]
        L.si_test_batch_plan.argtypes = [vp] + [i32] * 8 + [dbl] + [vp] * 12
        L.si_test_sinv_dense.argtypes = [i32]
        L.si_test_batch_graphs.argtypes = [i32, C.POINTER(i64), C.POINTER(i64)]
        L.si_test_spmm_shapes.argtypes = [vp, vp, i32, i32, i32, C.POINTER(i32), C.POINTER(i32)]
        L.si_version.argtypes = []
        L.si_mpnn_param_count.argtypes = []
        L.si_test_fast_solve.argtypes = [vp, vp, i32, u64, C.POINTER(_Config), vp, vp,
                                         C.POINTER(_Report)]
        L.si_test_band_solve.argtypes = [vp, vp, dbl, vp, vp, i32, C.POINTER(i32)]
        L.si_test_rcm_order.argtypes = [i64, vp, vp, i32, vp]
        L.si_test_laplacian_storage.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.si_test_components_device.argtypes = [vp, i64, vp, vp, vp, vp]
        L.si_test_rcm_order_device.argtypes = [vp, i64, vp, vp, vp]
        L.si_test_output_stage.argtypes = [i32]
        L.si_test_knn_path.argtypes = [i32, C.POINTER(i64)]
        L.si_profile_spmm_begin.argtypes = [vp]
        L.si_profile_spmm_end.argtypes = [vp, C.POINTER(_SpmmProfile)]
        L.si_profile_gather_probe.argtypes = [vp, vp, i32, i32, C.POINTER(dbl)]
        for name in EXPORTED:
            if name not in ("si_version", "si_last_error", "si_config_default",
                            "si_context_destroy", "si_laplacian_destroy", "si_mpnn_destroy",
                            "si_mpnn_param_count", "si_graph_destroy"):
                getattr(L, name).restype = C.c_int
        L.si_mpnn_param_count.restype = C.c_int64
        L.si_mpnn_destroy.restype = None
        L.si_mpnn_destroy.argtypes = [C.c_void_p]
        L.si_mpnn_create.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
        L.si_mpnn_get_params.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        L.si_mpnn_set_params.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        L.si_mpnn_features.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
        L.si_mpnn_forward.argtypes = [C.c_void_p] * 5
        L.si_mpnn_gradient.argtypes = [C.c_void_p] * 7
        L.si_mpnn_train_step.argtypes = [C.c_void_p] * 5 + [C.c_float, C.c_void_p]
        L.si_mpnn_warm_start.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int32,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.si_knn_exact.argtypes = [vp, vp, i32, i64, i64, i64, vp, vp]
        L.si_smooth_knn_dist.argtypes = [vp, vp, i64, i64, vp, vp]
        L.si_membership_strengths.argtypes = [vp, vp, vp, vp, vp, i64, i64, vp]
        L.si_fuzzy_union.argtypes = [vp, vp, vp, i64, i64, i32, C.POINTER(vp)]
        L.si_fuzzy_graph_build.argtypes = [vp, vp, i32, i64, i64, i64, i32, C.POINTER(vp)]
        L.si_graph_info.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
        L.si_graph_copy.argtypes = [vp, vp, vp, vp, vp]
        L.si_graph_destroy.restype = None
        L.si_graph_destroy.argtypes = [vp]
        _lib = L
    return _lib


def _ptr(a) -> Optional[int]:
    """Address of a numpy array (host) or a torch tensor (device or host)."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


class Context:
    """si_context: one HIP device and one stream."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        self._check(lib().si_context_create(device, C.byref(self._h)))

    def _check(self, st: int):
        if st != SI_OK:
            msg = lib().si_last_error(self._h).decode() if self._h else ""
            raise _ERRORS.get(st, SpectralError)(msg or f"status {st}")

    def close(self):
        if self._h:
            lib().si_context_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h


class Laplacian:
    """Device-resident L = I - D^-1/2 W D^-1/2 (src/laplacian.rs:29-88)."""

    def __init__(self, ctx: Context, indptr, indices, data, mode: int = 0, on_device=False,
                 reorder: Optional[bool] = None):
        """reorder: True stores the operator in a locality order (the clustering
        order for a connected graph of >= 65536 nodes, reverse Cuthill-McKee
        otherwise), "rcm" in reverse Cuthill-McKee order, False in the input
        order; None = the library default (a locality order for n >= 4096).
        Vectors always use the caller's order."""
        self.ctx = ctx
        self._h = C.c_void_p()
        if on_device:
            n = indptr.numel() - 1
            self._keep = (indptr, indices, data)
        else:
            indptr = np.ascontiguousarray(indptr, np.int64)
            indices = np.ascontiguousarray(indices, np.int32)
            data = np.ascontiguousarray(data, np.float32)
            n = indptr.size - 1
            self._keep = (indptr, indices, data)
        if reorder is None:
            ctx._check(lib().si_laplacian_build(ctx.handle, n, _ptr(indptr), _ptr(indices),
                                                _ptr(data), int(on_device), int(mode),
                                                C.byref(self._h)))
        else:
            flags = SI_BUILD_RCM_ORDER if reorder == "rcm" else (SI_BUILD_LOCALITY_ORDER if reorder else 0)
            ctx._check(lib().si_laplacian_build_ex(ctx.handle, n, _ptr(indptr), _ptr(indices),
                                                   _ptr(data), int(on_device), int(mode),
                                                   flags, C.byref(self._h)))
        n_, nnz, nl, ex = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        lib().si_laplacian_info(self._h, C.byref(n_), C.byref(nnz), C.byref(nl), C.byref(ex))
        self.n, self.nnz, self.n_long, self.degree_exact = n_.value, nnz.value, nl.value, bool(ex.value)

    def close(self):
        if self._h:
            lib().si_laplacian_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def export(self):
        """(degrees, sqrt_deg, (l_indptr, l_indices, l_data)) in the reference's CSR layout."""
        n = self.n
        deg = np.empty(n)
        sq = np.empty(n)
        ip = np.empty(n + 1, np.int64)
        ix = np.empty(self.nnz + n, np.int32)
        d = np.empty(self.nnz + n)
        m = C.c_int64()
        self.ctx._check(lib().si_laplacian_export(self.ctx.handle, self._h, _ptr(deg), _ptr(sq),
                                                  _ptr(ip), _ptr(ix), _ptr(d), C.byref(m)))
        return deg, sq, (ip, ix[:m.value].copy(), d[:m.value].copy())

    def spmm(self, x, y, precision=SI_F64):
        """y = L x for device blocks (CsrOperator::apply, src/operator.rs:176-228)."""
        k = x.shape[1] if x.dim() == 2 else 1
        ld = x.stride(0) if x.dim() == 2 else 1
        self.ctx._check(lib().si_spmm(self.ctx.handle, self._h, precision, _ptr(x), _ptr(y), k, ld))
        return y


def operator_apply_host(lap: Laplacian, x: np.ndarray) -> np.ndarray:
    """LinearOperator::apply on host arrays (src/operator.rs:6-12): y = L x, FP64."""
    x = np.ascontiguousarray(x, np.float64)
    xx = x[:, None] if x.ndim == 1 else x
    y = np.empty_like(xx)
    lap.ctx._check(lib().si_operator_apply_host(lap.ctx.handle, lap.handle, _ptr(xx), _ptr(y),
                                                xx.shape[1]))
    return y[:, 0] if x.ndim == 1 else y


def spmv_csr(ctx: Context, indptr, indices, data, x) -> np.ndarray:
    """spmv_csr (src/operator.rs:16-57) on the device, bit-identical to the
    reference's scalar loop: y = A x for a general CSR matrix."""
    ip = np.ascontiguousarray(indptr, np.uint64)
    ix = np.ascontiguousarray(indices, np.uint64)
    d = np.ascontiguousarray(data, np.float64)
    xv = np.ascontiguousarray(x, np.float64)
    y = np.empty(ip.size - 1, np.float64)
    ctx._check(lib().si_spmv_csr(ctx.handle, ip.size - 1, _ptr(ip), _ptr(ix), _ptr(d), d.size,
                                 _ptr(xv), xv.size, _ptr(y)))
    return y


def chebyshev_filter(lap: Laplacian, r, a=0.1, b=2.0, degree=8, out=None):
    import torch
    out = torch.empty_like(r) if out is None else out
    lap.ctx._check(lib().si_chebyshev_filter(lap.ctx.handle, lap.handle, _ptr(r), r.shape[1], a, b,
                                             degree, _ptr(out)))
    return out


def max_eigenpair_residual(lap: Laplacian, evals, evecs) -> float:
    ev = np.ascontiguousarray(evals, np.float64)
    out = C.c_double()
    lap.ctx._check(lib().si_max_eigenpair_residual(lap.ctx.handle, lap.handle, _ptr(ev),
                                                   _ptr(evecs), evecs.shape[1], C.byref(out)))
    return out.value


def max_eigenpair_residual_noalloc(lap: Laplacian, evals, evecs, k: Optional[int] = None) -> float:
    """The same residual through the entry point that allocates no device
    memory (its scratch is a static device array; overlapping calls race)."""
    ev = np.ascontiguousarray(evals, np.float64)
    out = C.c_double()
    kk = evecs.shape[1] if k is None else k
    lap.ctx._check(lib().si_max_eigenpair_residual_noalloc(lap.ctx.handle, lap.handle, _ptr(ev),
                                                           _ptr(evecs), kk, C.byref(out)))
    return out.value


def _evecs(lap, k):
    import torch
    return torch.empty((lap.n, k), dtype=torch.float64, device="cuda")


def lobpcg_solve(lap: Laplacian, n_components: int, seed: int, regularize=False, tol=1e-5):
    """lobpcg_solve (src/solvers/lobpcg.rs:227-465), device FP64."""
    k = n_components + 1
    ev = np.empty(k)
    vec = _evecs(lap, k)
    rep = _Report()
    rep.struct_size = C.sizeof(_Report)
    lap.ctx._check(lib().si_lobpcg_solve(lap.ctx.handle, lap.handle, n_components, seed,
                                         int(regularize), tol, _ptr(ev), _ptr(vec), C.byref(rep)))
    return ev, vec, Report._from(rep, k)


def lobpcg_sinv_solve(lap: Laplacian, n_components: int, seed: int):
    """lobpcg_sinv_solve (src/solvers/sinv.rs:106-189): shift-and-invert LOBPCG
    on (L + 1e-4 I)^-1, device FP64 (banded Cholesky).  Raises
    ConvergenceFailure for the reference's None."""
    k = n_components + 1
    ev = np.empty(k)
    vec = _evecs(lap, k)
    rep = _Report()
    rep.struct_size = C.sizeof(_Report)
    lap.ctx._check(lib().si_lobpcg_sinv_solve(lap.ctx.handle, lap.handle, n_components, seed,
                                              _ptr(ev), _ptr(vec), C.byref(rep)))
    return ev, vec, Report._from(rep, k)


def solve_eigenproblem(lap: Laplacian, n_components: int, seed: int,
                       config: Optional[SpectralInitConfig] = None):
    """solve_eigenproblem (src/solvers/mod.rs:70-224)."""
    k = n_components + 1
    cfg = (config or SpectralInitConfig())._c()
    ev = np.empty(k)
    vec = _evecs(lap, k)
    rep = _Report()
    rep.struct_size = C.sizeof(_Report)
    lap.ctx._check(lib().si_solve_eigenproblem(lap.ctx.handle, lap.handle, n_components, seed,
                                               C.byref(cfg), _ptr(ev), _ptr(vec), C.byref(rep)))
    return ev, vec, Report._from(rep, k)


def fast_solve(lap: Laplacian, n_components: int, seed: int,
               config: Optional[SpectralInitConfig] = None):
    """The FP32-bulk Chebyshev-filtered block solver alone (level 1)."""
    k = n_components + 1
    cfg = (config or SpectralInitConfig(tol=1e-6))._c()
    ev = np.empty(k)
    vec = _evecs(lap, k)
    rep = _Report()
    rep.struct_size = C.sizeof(_Report)
    lap.ctx._check(lib().si_test_fast_solve(lap.ctx.handle, lap.handle, n_components, seed,
                                            C.byref(cfg), _ptr(ev), _ptr(vec), C.byref(rep)))
    return ev, vec, Report._from(rep, k)


def rsvd_solve(lap: Laplacian, n_components: int, seed: int, oversampling=-1, nbiter=2):
    """rsvd_solve (src/solvers/rsvd.rs:135-213)."""
    k = n_components + 1
    ev = np.empty(k)
    vec = _evecs(lap, k)
    rep = _Report()
    rep.struct_size = C.sizeof(_Report)
    lap.ctx._check(lib().si_rsvd_solve(lap.ctx.handle, lap.handle, n_components, seed,
                                       oversampling, nbiter, _ptr(ev), _ptr(vec), C.byref(rep)))
    return ev, vec, Report._from(rep, k)


def dense_evd(lap: Laplacian, k: int):
    """dense_evd (src/solvers/dense.rs:9-46) on the device, n <= 4096, k <= 64."""
    ev = np.empty(k)
    vec = _evecs(lap, k)
    lap.ctx._check(lib().si_dense_evd(lap.ctx.handle, lap.handle, k, _ptr(ev), _ptr(vec)))
    return ev, vec


def spectral_init(graph, n_components: int, seed: int, data=None,
                  config: Optional[SpectralInitConfig] = None, ctx: Optional[Context] = None,
                  return_report: bool = False):
    """spectral_init (src/lib.rs:198-258).

    graph: (indptr, indices, data) CSR triple or a scipy.sparse CSR matrix
    (f32 weights, u32/i32 columns).  Returns an (n, n_components) float32 array.
    """
    if hasattr(graph, "indptr"):
        indptr, indices, weights = graph.indptr, graph.indices, graph.data
    else:
        indptr, indices, weights = graph
    indptr = np.ascontiguousarray(indptr, np.int64)
    indices = np.ascontiguousarray(indices, np.int32)
    weights = np.ascontiguousarray(weights, np.float32)
    n = indptr.size - 1
    own = ctx is None
    ctx = ctx or Context(0)
    cfg = (config or SpectralInitConfig())._c()
    raw = None if data is None else np.ascontiguousarray(data, np.float32)
    out = np.empty((max(n, 0), n_components), np.float32)
    rep = _Report()
    rep.struct_size = C.sizeof(_Report)
    try:
        ctx._check(lib().si_spectral_init(ctx.handle, n, _ptr(indptr), _ptr(indices), _ptr(weights),
                                          n_components, seed, _ptr(raw),
                                          0 if raw is None else raw.shape[1], C.byref(cfg),
                                          _ptr(out), C.byref(rep)))
    finally:
        if own:
            ctx.close()
    return (out, Report._from(rep, n_components + 1)) if return_report else out


# ---------------------------------------------------------------- kernel seams (tests)
def gram(ctx: Context, a, b, precision=None) -> np.ndarray:
    import torch
    prec = SI_F32 if a.dtype == torch.float32 else SI_F64
    p, q = a.shape[1], b.shape[1]
    g = np.empty((p, q))
    ctx._check(lib().si_test_gram(ctx.handle, prec, a.shape[0], _ptr(a), a.stride(0), p, _ptr(b),
                                  b.stride(0), q, _ptr(g)))
    return g


def block_mul(ctx: Context, a, c: np.ndarray, out):
    import torch
    prec = SI_F32 if a.dtype == torch.float32 else SI_F64
    c = np.ascontiguousarray(c, np.float64)
    ctx._check(lib().si_test_block_mul(ctx.handle, prec, a.shape[0], _ptr(a), a.stride(0),
                                       a.shape[1], _ptr(c), c.shape[1], _ptr(out), out.stride(0)))
    return out


def small_sym_eig(ctx: Context, a: np.ndarray):
    a = np.ascontiguousarray(a, np.float64)
    n = a.shape[0]
    w, v = np.empty(n), np.empty((n, n))
    ctx._check(lib().si_test_small_sym_eig(ctx.handle, n, _ptr(a), _ptr(w), _ptr(v)))
    return w, v


def small_gen_eig(ctx: Context, a: np.ndarray, b: np.ndarray):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    n = a.shape[0]
    w, c = np.empty(n), np.empty((n, n))
    ctx._check(lib().si_test_small_gen_eig(ctx.handle, n, _ptr(a), _ptr(b), _ptr(w), _ptr(c)))
    return w, c


def fill_normal(ctx: Context, x, seed: int, stream: int):
    ctx._check(lib().si_test_fill_normal(ctx.handle, _ptr(x), x.shape[0], x.shape[1], seed, stream))


def fill_normal_f32(ctx: Context, x, seed: int, stream: int):
    """Testing seam: the FP32 start-block stream into a device float32 [n, k] block."""
    ctx._check(lib().si_test_fill_normal_f32(ctx.handle, _ptr(x), x.shape[0], x.shape[1], seed,
                                             stream))
    return x


def staged_roundtrip(ctx: Context, data: np.ndarray) -> np.ndarray:
    """Testing seam: host bytes up to the device and back through the staged copies."""
    src = np.ascontiguousarray(data)
    out = np.empty_like(src)
    ctx._check(lib().si_test_staged_roundtrip(ctx.handle, src.ctypes.data, out.ctypes.data,
                                              src.nbytes))
    return out


class SpmmProfiler:
    """Event-timed SpMM launches on the context's stream (bytes are algorithmic)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.result = None

    def __enter__(self):
        self.ctx._check(lib().si_profile_spmm_begin(self.ctx.handle))
        return self

    def __exit__(self, *exc):
        p = _SpmmProfile()
        self.ctx._check(lib().si_profile_spmm_end(self.ctx.handle, C.byref(p)))
        self.result = dict(launches=p.launches, bytes=p.bytes, ms=p.ms,
                           launches_f32=p.launches_f32, bytes_f32=p.bytes_f32, ms_f32=p.ms_f32)
        return False


def gather_probe(lap: "Laplacian", mode: int, reps: int = 20) -> float:
    """Microseconds per launch of the gather probe (mode 0: the SpMM's own
    64-byte row gathers along the column stream; 1: random rows; 2: random rows
    in a 2 MiB footprint)."""
    us = C.c_double()
    lap.ctx._check(lib().si_profile_gather_probe(lap.ctx.handle, lap.handle, int(mode), int(reps),
                                                 C.byref(us)))
    return us.value


def batch_graphs(enable: Optional[bool] = None):
    """Testing seam: switch hipGraph capture of batched component solves
    (None leaves it) and return (graphs captured, graph launches) of the last
    batched solve."""
    cap, lau = C.c_int64(0), C.c_int64(0)
    lib().si_test_batch_graphs(-1 if enable is None else int(bool(enable)), C.byref(cap),
                               C.byref(lau))
    return cap.value, lau.value


def sinv_dense(enable: bool) -> None:
    """Testing seam: factorise the shift-and-invert operator densely whatever
    its band (True), or choose automatically (False)."""
    lib().si_test_sinv_dense(C.c_int32(1 if enable else 0))


def batch_plan(ctx: Context, mode: int, update: int, state: dict, status, theta, norms2, k: int,
               m: int, cap: int, steps: int, max_rounds: int, tol: float):
    """Testing seam: one call of the batched solver's device round planner.
    state holds a, best, worst (float64 [nseg]), stall, done (int32 [nseg]),
    ctl (int32 [4]) and, for mode 1, thk (float64 [nseg, m]); it is updated in
    place.  Returns (deg [nseg], coef [steps, nseg + 1, 3])."""
    a = np.ascontiguousarray(state["a"], np.float64)
    nseg = a.size
    arrs = {"a": a}
    for key, dt in (("best", np.float64), ("worst", np.float64), ("stall", np.int32),
                    ("done", np.int32), ("ctl", np.int32)):
        arrs[key] = np.ascontiguousarray(state[key], dt).copy()
    arrs["a"] = a.copy()
    thk = None
    if mode == 1:
        thk = np.ascontiguousarray(state["thk"], np.float64).copy()
    status = np.ascontiguousarray(status, np.int32)
    theta = np.ascontiguousarray(theta, np.float64)
    norms2 = np.ascontiguousarray(norms2, np.float64)
    assert theta.size == nseg * m and norms2.size == nseg * 2 * k and status.size == nseg
    deg = np.zeros(nseg, np.int32)
    coef = np.zeros((steps, nseg + 1, 3))
    P = lambda x: C.c_void_p(x.ctypes.data) if x is not None else C.c_void_p(0)
    ctx._check(lib().si_test_batch_plan(
        ctx.handle, C.c_int32(mode), C.c_int32(update), C.c_int32(nseg), C.c_int32(k),
        C.c_int32(m), C.c_int32(cap), C.c_int32(steps), C.c_int32(max_rounds), C.c_double(tol),
        P(arrs["a"]), P(arrs["best"]), P(arrs["worst"]), P(arrs["stall"]), P(arrs["done"]),
        P(thk), P(status), P(theta), P(norms2), P(arrs["ctl"]), P(deg), P(coef)))
    for key in arrs:
        state[key] = arrs[key]
    if thk is not None:
        state["thk"] = thk
    return deg, coef


def spmm_shapes(lap: "Laplacian", shape16: int = 0, shape32: int = 0, autotune: bool = False):
    """Testing seam: set (or autotune) the FP32 16/32-column SpMM shapes of a
    Laplacian; returns the shapes in use."""
    a, b = C.c_int32(0), C.c_int32(0)
    lap.ctx._check(lib().si_test_spmm_shapes(lap.ctx.handle, lap.handle, int(shape16),
                                             int(shape32), int(bool(autotune)), C.byref(a),
                                             C.byref(b)))
    return a.value, b.value


def band_solve(lap: "Laplacian", x, shift: float = 1e-4):
    """Testing seam: (L + shift I)^-1 x for a device [n x k] f64 block through the
    banded Cholesky of the shift-and-invert level; returns (y, bandwidth)."""
    import torch
    x = x.contiguous()
    k = x.shape[1] if x.dim() == 2 else 1
    y = torch.empty_like(x)
    bw = C.c_int32(0)
    lap.ctx._check(lib().si_test_band_solve(lap.ctx.handle, lap.handle, float(shift), _ptr(x),
                                            _ptr(y), k, C.byref(bw)))
    return y, bw.value


# ---------------------------------------------------------------- learned warm start
class WarmStartNet:
    """The learned warm-start network (include/spectral_mi355x_warmstart.h)."""

    def __init__(self, ctx: Context, seed: int = 0, params: Optional[np.ndarray] = None):
        self.ctx = ctx
        h = C.c_void_p()
        ctx._check(lib().si_mpnn_create(ctx.handle, seed, C.byref(h)))
        self._h = h
        if params is not None:
            self.set_params(params)

    @staticmethod
    def param_count() -> int:
        return int(lib().si_mpnn_param_count())

    def params(self) -> np.ndarray:
        out = np.empty(self.param_count(), np.float32)
        self.ctx._check(lib().si_mpnn_get_params(self.ctx.handle, self._h, out.ctypes.data, out.size))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        self.ctx._check(lib().si_mpnn_set_params(self.ctx.handle, self._h, p.ctypes.data, p.size))

    def close(self):
        if getattr(self, "_h", None):
            lib().si_mpnn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def features(lap: "Laplacian", seed: int = 0):
        import torch
        f = torch.empty((lap.n, 16), dtype=torch.float32, device="cuda")
        lap.ctx._check(lib().si_mpnn_features(lap.ctx.handle, lap.handle, seed, f.data_ptr()))
        return f

    def forward(self, lap: "Laplacian", f):
        import torch
        x = torch.empty((lap.n, 8), dtype=torch.float32, device="cuda")
        self.ctx._check(lib().si_mpnn_forward(self.ctx.handle, self._h, lap.handle, f.data_ptr(),
                                              x.data_ptr()))
        return x

    def gradient(self, lap: "Laplacian", f, v):
        loss = C.c_double(0.0)
        g = np.empty(self.param_count(), np.float64)
        self.ctx._check(lib().si_mpnn_gradient(self.ctx.handle, self._h, lap.handle, f.data_ptr(),
                                               v.data_ptr(), C.byref(loss), g.ctypes.data))
        return loss.value, g

    def train_step(self, lap: "Laplacian", f, v, lr: float) -> float:
        loss = C.c_double(0.0)
        self.ctx._check(lib().si_mpnn_train_step(self.ctx.handle, self._h, lap.handle, f.data_ptr(),
                                                 v.data_ptr(), float(lr), C.byref(loss)))
        return loss.value

    def warm_start(self, lap: "Laplacian", seed: int = 0, n_components: int = 2):
        import torch
        blk = torch.empty((lap.n, n_components + 1), dtype=torch.float64, device="cuda")
        used = C.c_int32(0)
        rn, rr = C.c_double(0.0), C.c_double(0.0)
        self.ctx._check(lib().si_mpnn_warm_start(self.ctx.handle, self._h, lap.handle, seed,
                                                 n_components, blk.data_ptr(), C.byref(used),
                                                 C.byref(rn), C.byref(rr)))
        return blk, bool(used.value), rn.value, rr.value


# ------------------------------------------------------------------ fuzzy k-NN graph
def _points(x):
    x = np.ascontiguousarray(x)
    if x.dtype == np.float64:
        return x, SI_F64
    return np.ascontiguousarray(x, np.float32), SI_F32


def knn_exact(ctx: Context, x, k: int):
    """Exact k-NN on the device, self excluded (tests/generate_fixtures.py:70-93):
    (indices int32 [n, k], distances float32 [n, k])."""
    x, prec = _points(x)
    n, d = x.shape
    idx = np.empty((n, k), np.int32)
    dist = np.empty((n, k), np.float32)
    ctx._check(lib().si_knn_exact(ctx.handle, x.ctypes.data, prec, n, d, k, idx.ctypes.data,
                                  dist.ctypes.data))
    return idx, dist


def smooth_knn_dist(ctx: Context, dist):
    """umap smooth_knn_dist on the device: (sigmas, rhos) float32 [n]."""
    dist = np.ascontiguousarray(dist, np.float32)
    n, k = dist.shape
    sig = np.empty(n, np.float32)
    rho = np.empty(n, np.float32)
    ctx._check(lib().si_smooth_knn_dist(ctx.handle, dist.ctypes.data, n, k, sig.ctypes.data,
                                        rho.ctypes.data))
    return sig, rho


def membership_strengths(ctx: Context, idx, dist, sigmas, rhos):
    """umap compute_membership_strengths on the device: float32 [n, k]."""
    idx = np.ascontiguousarray(idx, np.int32)
    dist = np.ascontiguousarray(dist, np.float32)
    sigmas = np.ascontiguousarray(sigmas, np.float32)
    rhos = np.ascontiguousarray(rhos, np.float32)
    n, k = idx.shape
    vals = np.empty((n, k), np.float32)
    ctx._check(lib().si_membership_strengths(ctx.handle, idx.ctypes.data, dist.ctypes.data,
                                             sigmas.ctypes.data, rhos.ctypes.data, n, k,
                                             vals.ctypes.data))
    return vals


class Graph:
    """si_graph: a CSR graph on the device (int64 indptr, int32 indices, float32)."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self._h = handle
        n, nnz = C.c_int64(0), C.c_int64(0)
        ctx._check(lib().si_graph_info(self._h, C.byref(n), C.byref(nnz)))
        self.n, self.nnz = n.value, nnz.value

    def to_host(self):
        ip = np.empty(self.n + 1, np.int64)
        ix = np.empty(max(self.nnz, 1), np.int32)
        w = np.empty(max(self.nnz, 1), np.float32)
        self.ctx._check(lib().si_graph_copy(self.ctx.handle, self._h, ip.ctypes.data,
                                            ix.ctypes.data, w.ctypes.data))
        return ip, ix[:self.nnz], w[:self.nnz]

    def close(self):
        if self._h:
            lib().si_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fuzzy_union(ctx: Context, idx, vals, n_epochs: int = -1) -> Graph:
    """W = A + A^T - A o A^T on the device, pruned below max/n_epochs
    (n_epochs -1: 500 for n <= 10000 else 200; 0: no pruning)."""
    idx = np.ascontiguousarray(idx, np.int32)
    vals = np.ascontiguousarray(vals, np.float32)
    n, k = idx.shape
    h = C.c_void_p()
    ctx._check(lib().si_fuzzy_union(ctx.handle, idx.ctypes.data, vals.ctypes.data, n, k,
                                    int(n_epochs), C.byref(h)))
    return Graph(ctx, h)


def fuzzy_graph(ctx: Context, x, k: int = 15, n_epochs: int = -1) -> Graph:
    """Steps 1-5a of the fuzzy k-NN graph on the device, intermediates resident."""
    x, prec = _points(x)
    n, d = x.shape
    h = C.c_void_p()
    ctx._check(lib().si_fuzzy_graph_build(ctx.handle, x.ctypes.data, prec, n, d, k,
                                          int(n_epochs), C.byref(h)))
    return Graph(ctx, h)


ORDER_INPUT, ORDER_RCM, ORDER_CLUSTER = 0, 1, 2


def laplacian_storage(lap: "Laplacian"):
    """Testing seam: the operator as stored on the device -- (rowptr, col, a32,
    perm, order) in the storage node order, perm[new] = old, order one of
    ORDER_INPUT, ORDER_RCM, ORDER_CLUSTER."""
    n, nnz = lap.n, lap.nnz
    rp = np.empty(n + 1, np.int32)
    col = np.empty(nnz, np.int32)
    a32 = np.empty(nnz, np.float32)
    perm = np.empty(n, np.int32)
    order = C.c_int32()
    lap.ctx._check(lib().si_test_laplacian_storage(lap.ctx.handle, lap._h, rp.ctypes.data,
                                                   col.ctypes.data, a32.ctypes.data,
                                                   perm.ctypes.data, C.byref(order)))
    return rp, col, a32, perm, order.value


def set_output_stage(device: bool) -> None:
    """Testing seam: form spectral_init's output on the device (default) or on
    the host (the same values)."""
    st = lib().si_test_output_stage(1 if device else 0)
    if st != SI_OK:
        raise _ERRORS.get(st, SpectralError)(f"status {st}")


def components_device(ctx: Context, indptr, indices):
    """Testing seam: (labels, count) of a symmetric pattern, computed on the device."""
    ip = np.ascontiguousarray(indptr, np.int64)
    ix = np.ascontiguousarray(indices, np.int32)
    n = ip.size - 1
    lab = np.empty(n, np.int32)
    nc = C.c_int64()
    ctx._check(lib().si_test_components_device(ctx.handle, n, ip.ctypes.data,
                                               ix.ctypes.data if ix.size else None,
                                               lab.ctypes.data, C.byref(nc)))
    return lab, nc.value


def rcm_order_device(ctx: Context, indptr, indices) -> np.ndarray:
    """Testing seam: the locality order of a symmetric pattern computed on the
    device, perm[new] = old."""
    ip = np.ascontiguousarray(indptr, np.int64)
    ix = np.ascontiguousarray(indices, np.int32)
    n = ip.size - 1
    perm = np.empty(n, np.int32)
    ctx._check(lib().si_test_rcm_order_device(ctx.handle, n, ip.ctypes.data,
                                              ix.ctypes.data if ix.size else None,
                                              perm.ctypes.data))
    return perm


def laplacian_order(lap: "Laplacian") -> int:
    """Testing seam: which node order the operator is stored in (ORDER_INPUT,
    ORDER_RCM or ORDER_CLUSTER)."""
    order = C.c_int32()
    lap.ctx._check(lib().si_test_laplacian_storage(lap.ctx.handle, lap._h, None, None, None, None,
                                                   C.byref(order)))
    return order.value


def rcm_order(indptr, indices, threads: int = 0) -> np.ndarray:
    """Testing seam (host only): the locality order of si_laplacian_build,
    perm[new] = old."""
    ip = np.ascontiguousarray(indptr, np.int64)
    ix = np.ascontiguousarray(indices, np.int32)
    n = ip.size - 1
    perm = np.empty(n, np.int32)
    st = lib().si_test_rcm_order(n, ip.ctypes.data, ix.ctypes.data if ix.size else None,
                                 int(threads), perm.ctypes.data)
    if st != SI_OK:
        raise _ERRORS.get(st, SpectralError)(f"status {st}")
    return perm


def knn_path(mode: int = -1) -> int:
    """Testing seam: select the exact k-NN arithmetic of later calls (0
    automatic, 1 the direct FP64 kernel only, 2 the matrix-core filter with
    every query re-run directly, 3 the filter on the FP32 matrix cores only; -1
    keeps the mode) and return the number of queries the last filtered call
    re-ran with the direct kernel."""
    out = C.c_int64(0)
    st = lib().si_test_knn_path(int(mode), C.byref(out))
    if st != SI_OK:
        raise _ERRORS.get(st, SpectralError)(f"status {st}")
    return int(out.value)
