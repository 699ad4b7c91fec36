"""Host code under the address and undefined-behaviour sanitizers (CPU only).

* The layout harness (computational-mathematics_amd/check/layout_check.cpp)
  runs the engine's host-side index code -- CSR validation, SpMV row blocks,
  row bins, transposes, symmetric expansion, level schedules, the slot-major
  and diagonal indexes -- over random and edge-case matrices, checking each
  result, in a sanitized build.
* The host library's own tests (factor producers, Matrix Market ingest,
  generators) run in a subprocess against the sanitized build of
  libcmhost.so, with libasan preloaded into the interpreter.
Both binaries are built by `make -C computational-mathematics_amd` (target
`sanitize`).
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "computational-mathematics_amd")


def _ensure_built():
    if not (os.path.exists(os.path.join(PKG, "bin", "layout_check"))
            and os.path.exists(os.path.join(PKG, "lib", "asan", "libcmhost.so"))):
        subprocess.run(["make", "-C", PKG, "sanitize"], check=True, capture_output=True)


def _runtime(name):
    out = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


def test_layout_harness_under_sanitizers():
    _ensure_built()
    r = subprocess.run([os.path.join(PKG, "bin", "layout_check")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "all layout checks passed" in r.stdout


def test_host_library_tests_under_sanitizers():
    _ensure_built()
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("libasan runtime not found")
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"  # the interpreter itself keeps memory at exit
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["CMAMD_HOST_LIB"] = os.path.join(PKG, "lib", "asan", "libcmhost.so")
    env["OMP_NUM_THREADS"] = "2"
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_host.py"), os.path.join(ROOT, "tests", "test_factors.py")],
                       capture_output=True, text=True, env=env, timeout=900, cwd=ROOT)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "passed" in r.stdout
