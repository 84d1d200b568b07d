"""The C-ABI library on the CPU: it loads, exports every declared symbol, and
its host-side logic (initial conditions, Yoshida coefficients, argument
checks) matches the oracle.  No kernel is launched here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import nbody_amd
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "nbody.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^NB_API\s+[\w\s\*]+?\b(nb_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = nbody_amd.load_library()
    names = declared_symbols()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), name
    assert sorted(nbody_amd.EXPORTS) == names


def test_version_and_error_text():
    L = nbody_amd.load_library()
    assert b"gfx950" in L.nb_version()
    assert L.nb_yoshida(2, None, None, None) == 1   # NB_ERR_ARG
    assert b"nb_yoshida" in L.nb_last_error()


@pytest.mark.parametrize("kind,n,seed", [("sphere", 2048, 42), ("plummer", 4096, 7), ("plummer", 1, 3),
                                         ("sphere", 0, 1)])
def test_initial_conditions_match_oracle_bitwise(orc, kind, n, seed):
    m, pos, vel, _ = nbody_amd.init_system(kind, n, seed)
    if kind == "sphere":
        mo, po, vo = orc.ic_sphere(n, seed)
    else:
        mo, po, vo = orc.ic_plummer(n, seed, 1.0, 20.0)
    assert np.array_equal(m, mo) and np.array_equal(pos, po) and np.array_equal(vel, vo)


@pytest.mark.parametrize("n", [2, 3, 64, 2049])
def test_ring_matches_oracle_bitwise(orc, n):
    m, pos, vel, f = nbody_amd.init_system("ring", n, 0, [1.0, 2 * np.pi, 1e5])
    mo, po, vo, fo = orc.ic_ring(n, 1.0, 2 * np.pi, 1e5)
    for a, b in ((m, mo), (pos, po), (vel, vo), (f, fo)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("order", [1, 2, 3, 4, 6, 8])
def test_yoshida_matches_oracle(orc, order):
    c, d, lf = nbody_amd.yoshida(order)
    co, do, lfo = orc.yoshida(order)
    assert lf == lfo and np.array_equal(c, co) and np.array_equal(d, do)


def test_argument_errors():
    L = nbody_amd.load_library()
    m = np.zeros(4)
    planes = (C.POINTER(C.c_double) * 3)()
    # unknown IC kind, null outputs, Plummer scale <= 0
    assert L.nb_init_system(7, 4, 1, None, m.ctypes.data_as(C.POINTER(C.c_double)), planes, planes, None) == 1
    assert L.nb_init_system(1, 4, 1, None, None, None, None, None) == 1
    pos, vel = np.zeros((3, 4)), np.zeros((3, 4))
    bad = np.array([-1.0, 20.0])
    with pytest.raises(nbody_amd.NBodyError):
        nbody_amd.init_system("plummer", 4, 1, bad)
    # null context everywhere
    assert L.nb_set_state(None, 0, None, None, None) == 1
    assert L.nb_step(None, 1, 1e-3, None, None, 2, 0, 0.0) == 1
    assert L.nb_energy(None, 0.0, None, None) == 1
    assert L.nb_step_euler(None, 1, 1e-3, 0, 0.0) == 1
    assert L.nb_run_records(None, 0, 0.0, 1.0, None, None, None, None, None, 0, 1, 1e-3, None, None, 2, None) == 1
    assert L.nb_step_rk(None, 1, 1e-3, None, None, 4, 0, 0.0) == 1
    assert L.nb_runge_kutta(4, None, None, None) == 1
    assert L.nb_kernel_info_read(None, None) == 1
    assert L.nb_destroy(None) == 0


def test_cli_refuses_run_types_it_does_not_provide():
    # the reference's cache/scaling/MPI/hybrid run types end with a message
    # before any device is touched
    cli = os.path.join(ROOT, "n-body-ring-solver_amd", "bin", "nbody")
    if not os.path.exists(cli):
        pytest.skip("CLI not built")
    for t in ("omp-weak", "hybrid", "mpi-bandwidth"):
        r = subprocess.run([cli, "-t", t, "-N", "16"], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and "not provided" in r.stdout
    r = subprocess.run([cli, "-t", "bogus", "-i", "bogus", "-S", "x", "-p", "x", "-V", "x", "-N", "16", "-device", "99"],
                       capture_output=True, text=True, timeout=60)
    for msg in ("Unknown Execution type: bogus", "Defaulting to omp", "Unknown Integration kind: bogus",
                "Defaulting to euler", "Unknown Solver type: x", "Unknown Data type: x", "Unknown Vectorization type: x"):
        assert msg in r.stdout, msg
    assert r.returncode != 0


def _gpu_present() -> bool:
    L = nbody_amd.load_library()
    h = C.c_void_p()
    ok = L.nb_create(C.byref(h), 0) == 0
    if ok:
        L.nb_destroy(h)
    return ok


def test_no_device_fails_loudly():
    if _gpu_present():
        pytest.skip("a GPU is present")
    with pytest.raises(nbody_amd.NBodyError) as e:
        nbody_amd.Solver(0)
    assert e.value.status == 2   # NB_ERR_DEVICE
    cli = os.path.join(ROOT, "n-body-ring-solver_amd", "bin", "nbody")
    if os.path.exists(cli):
        r = subprocess.run([cli, "-N", "16", "-s", "1"], capture_output=True, text=True, timeout=60)
        assert r.returncode != 0 and "nb_create" in r.stderr
