"""Shared test set-up.

Markers: `gpu` tests need a gfx950 device and call the solver through its C
ABI (n-body-ring-solver_amd/lib/libnbody_hip.so); everything else runs on the
CPU (the oracle, the golden fixtures, the host logic and the library's
exports).  The oracle (oracle/) is the checker only.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "n-body-ring-solver_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; calls the HIP library")


@pytest.fixture(scope="session")
def orc():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle import REF_SO, Reference, cpu_has_avx512
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref/libnbref.so not built (reference sources absent)")
    if not cpu_has_avx512():
        pytest.skip("host CPU lacks AVX-512; the reference build targets x86-64-v4")
    return Reference()


@pytest.fixture(scope="session")
def solver():
    import nbody_amd
    s = nbody_amd.Solver(0)
    yield s
    s.close()


def golden(name: str) -> dict:
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """max_i |a_i - b_i| / |b_i| over particles (columns) with |b_i| > 0."""
    nb = np.linalg.norm(b, axis=0)
    d = np.linalg.norm(a - b, axis=0)
    mask = nb > 0
    return float(np.max(d[mask] / nb[mask])) if mask.any() else 0.0


class DeviceArray:
    """A float64 array in device memory through the HIP runtime the solver
    library itself loaded (hipMalloc / hipMemcpy / hipFree), for the tests of
    the entry points that take device pointers."""

    def __init__(self, host: np.ndarray):
        import ctypes as C
        import nbody_amd
        nbody_amd.load_library()
        self._hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)   # the runtime the library links
        self._C = C
        self.host = np.ascontiguousarray(host, dtype=np.float64)
        self.ptr = C.c_void_p()
        assert self._hip.hipMalloc(C.byref(self.ptr), C.c_size_t(max(self.host.nbytes, 8))) == 0
        assert self._hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p), C.c_size_t(self.host.nbytes),
                                   1) == 0   # hipMemcpyHostToDevice

    @property
    def addr(self) -> int:
        return self.ptr.value

    def get(self) -> np.ndarray:
        out = np.empty_like(self.host)
        assert self._hip.hipDeviceSynchronize() == 0
        assert self._hip.hipMemcpy(out.ctypes.data_as(self._C.c_void_p), self.ptr, self._C.c_size_t(out.nbytes),
                                   2) == 0   # hipMemcpyDeviceToHost
        return out

    def free(self):
        if self.ptr:
            self._hip.hipFree(self.ptr)
            self.ptr = self._C.c_void_p()
