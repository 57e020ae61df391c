"""Sanitizer builds of the host-only locality order (CPU).

The threaded reverse Cuthill-McKee of si_laplacian_build (reorder.cpp: level
expansion on host threads with atomic-minimum parents) is built with g++ under
AddressSanitizer + UndefinedBehaviorSanitizer and under ThreadSanitizer
(spectral-init_amd/sanitize/Makefile) and run over random graphs whose BFS
levels take the threaded path, graphs with shuffled rows and repeated columns,
several components, a ring and a tiny graph; each order must be a permutation
and equal on 1 and 16 threads, and the sanitizers abort on any report.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "spectral-init_amd", "sanitize")
OUT = os.path.join(ROOT, "spectral-init_amd", "_build", "sanitize")


@pytest.fixture(scope="module")
def built():
    if not shutil.which("g++") or not os.path.exists("/opt/rocm/include/hip/hip_runtime.h"):
        pytest.skip("g++ or the HIP headers are missing")
    subprocess.run(["make", "-s", "-C", SAN], check=True, capture_output=True, text=True)
    return OUT


@pytest.mark.parametrize("flavour", ["asan", "tsan"])
def test_locality_order_under_sanitizers(built, flavour):
    exe = os.path.join(built, f"rcm_check_{flavour}")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all cases ok" in r.stdout
    assert "Sanitizer" not in r.stderr, r.stderr
