"""Sanitizer runs on the CPU: the oracle (every function, small and edge-case
inputs) and the library's host-only units (initial conditions, integrator
coefficients, snapshots, accuracy metrics, error state), each built with
-fsanitize=address,undefined and -fno-sanitize-recover, so any
out-of-bounds access, leak or undefined behaviour fails the test.  No GPU
code is built or run here."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "n-body-ring-solver_amd")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None,
                                reason="needs gcc/g++ with the sanitizer runtimes")


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, env=ENV, timeout=600, **kw)
    assert r.returncode == 0, f"{' '.join(cmd)} failed ({r.returncode}):\n{r.stdout[-4000:]}\n{r.stderr[-4000:]}"
    return r


def test_oracle_under_sanitizers():
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "sanitize"])
    r = _run([os.path.join(ROOT, "oracle", "_build", "sanitize_check")])
    assert "oracle sanitizer run: ok" in r.stdout
    assert "runtime error" not in r.stderr


def test_library_host_units_under_sanitizers(tmp_path):
    exe = str(tmp_path / "host_sanitize_check")
    srcs = [os.path.join(ROOT, "tests", "host_sanitize_check.cpp")] + [
        os.path.join(PKG, "csrc", f) for f in ("nbody_ic.cpp", "nbody_snapshot.cpp", "nbody_metrics.cpp", "nbody_status.cpp")]
    _run(["g++", "-std=c++17", "-O1", "-g", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
          "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(PKG, "csrc"),
          "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
          *srcs, "-o", exe, "-lm"])
    r = _run([exe, str(tmp_path)])
    assert "host sanitizer run: ok" in r.stdout
    assert "runtime error" not in r.stderr
