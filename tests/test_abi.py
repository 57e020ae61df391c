"""CPU checks of the C-ABI library: it loads without a GPU and exports every
entry point the public headers declare."""
import ctypes
import os
import re

import pytest

import spectral_mi355x as si

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "spectral_mi355x.h"),
           os.path.join(ROOT, "include", "spectral_mi355x_testing.h"),
           os.path.join(ROOT, "include", "spectral_mi355x_warmstart.h"),
           os.path.join(ROOT, "include", "spectral_mi355x_graph.h")]


def _declared():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(si_[a-z0-9_]+)\s*\(", text))
    return names


@pytest.fixture(scope="module")
def library():
    if not os.path.exists(si.library_path()):
        pytest.fail("native library not built (run spectral-init_amd/build.py)")
    return ctypes.CDLL(si.library_path())


def test_headers_declare_the_bound_symbols():
    assert _declared() == set(si.EXPORTED)


def test_library_exports_every_declared_symbol(library):
    missing = [n for n in _declared() if not hasattr(library, n)]
    assert not missing, missing


def test_every_binding_declares_its_argument_types():
    # without argtypes ctypes passes a Python int as a 32-bit C int, which
    # truncates a device address handed over as an integer
    lib = si.lib()
    undeclared = [n for n in si.EXPORTED if getattr(lib, n).argtypes is None]
    assert not undeclared, undeclared


def test_version_and_default_config_without_gpu(library):
    lib = si.lib()
    assert b"gfx950" in lib.si_version()
    c = si._Config()
    lib.si_config_default(ctypes.byref(c))
    # the reference defaults (PythonCompat, tol 1e-5, dense threshold 2000); degree 0 = automatic
    assert c.struct_size == ctypes.sizeof(si._Config)
    assert c.compute_mode == 0 and c.tol == 1e-5 and c.dense_threshold == 2000
    assert c.cheb_degree == 0
    assert c.batch_components == 1


def test_context_create_fails_cleanly_without_a_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    st = si.lib().si_context_create(0, ctypes.byref(h))
    assert st != si.SI_OK and not h.value


def test_product_does_not_link_the_oracle(library):
    # the shared library must not depend on the CPU oracle
    import subprocess
    out = subprocess.run(["ldd", si.library_path()], capture_output=True, text=True).stdout
    assert "oracle" not in out
    assert not hasattr(library, "or_spmm")


def test_knn_path_seam_without_gpu(library):
    """The k-NN path selector is host-only: modes -1..3 are accepted (and the
    mode persists until reset), anything else is SI_ERR_INVALID_ARGUMENT."""
    out = ctypes.c_int64(-7)
    fn = library.si_test_knn_path
    fn.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]
    fn.restype = ctypes.c_int
    try:
        for mode in (0, 1, 2, 3, -1):
            assert fn(mode, ctypes.byref(out)) == si.SI_OK
            assert out.value >= 0
        assert fn(4, ctypes.byref(out)) == si.SI_ERR_INVALID_ARGUMENT
        assert fn(-2, None) == si.SI_ERR_INVALID_ARGUMENT
        assert fn(-1, None) == si.SI_OK
    finally:
        fn(0, None)


def _rust_struct_fields(text, name):
    body = re.search(r"pub struct " + name + r" \{(.*?)\n\}", text, flags=re.S).group(1)
    return re.findall(r"pub (\w+):\s*([^,]+),", body)


def test_rust_shim_matches_the_c_abi():
    """rust/src/mi355x.rs (compile-unverified: no cargo here) declares only
    entry points the public header has, and its repr(C) structs list the C
    structs' fields in the same order with the same widths."""
    text = open(os.path.join(ROOT, "rust", "src", "mi355x.rs")).read()
    used = set(re.findall(r"pub fn (si_[a-z0-9_]+)\(", text))
    assert used and used <= _declared(), used - _declared()
    rust_size = {"u32": 4, "i32": 4, "i64": 8, "f64": 8, "*const f64": 8}
    for rname, cstruct in (("SiConfig", si._Config), ("SiReport", si._Report)):
        fields = _rust_struct_fields(text, rname)
        assert [f for f, _ in fields] == [f for f, _ in cstruct._fields_], rname
        for (f, ty), (_, cty) in zip(fields, cstruct._fields_):
            m = re.match(r"\[f64; (\d+)\]", ty.strip())
            size = 8 * int(m.group(1)) if m else rust_size[ty.strip()]
            assert size == ctypes.sizeof(cty), (rname, f)
