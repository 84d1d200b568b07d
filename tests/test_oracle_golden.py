"""The oracle against the reference's own outputs (CPU only).

Pins the oracle's force, energy and leapfrog restatements on the committed
golden fixtures (written by tests/golden/make_golden.py from the reference's
compiled CPU routines) and, where oracle/_ref is built, live against the
reference on fresh inputs.
"""
import numpy as np
import pytest

from conftest import golden


def scaled_err(a, b):
    """max |a - b| / max |b| over all components."""
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("n", [64, 257, 2049])
def test_ring_forces_match_reference_kernels(orc, n):
    g = golden(f"ring_{n}.npz")
    f = orc.accel(g["m"], g["pos"], 0.0) * g["m"]
    # the ring's analytic force (CPU/nbody_system.cpp:262-264) and both reference kernels
    assert scaled_err(f, g["force_analytic"]) < 1e-12
    assert scaled_err(f, g["force_avx2"]) < 1e-12
    assert scaled_err(f, g["force_avx512"]) < 1e-12


@pytest.mark.parametrize("n", [64, 300])
def test_jittered_forces_match_reference_kernels(orc, n):
    g = golden(f"jitter_{n}.npz")
    f = orc.accel(g["m"], g["pos"], 0.0) * g["m"]
    for key in ("force_avx2", "force_avx512"):
        per = np.linalg.norm(f - g[key], axis=0) / np.linalg.norm(g[key], axis=0)
        assert per.max() < 1e-12, key
    # the reduced (Newton's 3rd law) restatement too
    fr = orc.forces_reduced(g["m"], g["pos"], 0.0)
    assert scaled_err(fr, g["force_avx2"]) < 1e-12


@pytest.mark.parametrize("n", [64, 300])
def test_energies_match_reference(orc, n):
    g = golden(f"jitter_{n}.npz")
    T, U = orc.energy(g["m"], g["pos"], g["vel"], 0.0)
    assert abs(T - g["T"]) <= 1e-14 * abs(g["T"])
    assert abs(U - g["U_avx2"]) <= 1e-14 * abs(g["U_avx2"])
    assert abs(U - g["U_scalar"]) <= 1e-14 * abs(g["U_scalar"])


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_leapfrog_matches_reference_driver(orc, order):
    g = golden("jitter_64.npz")
    c, d, lf = orc.yoshida(order)
    assert np.array_equal(c, g[f"c{order}"]) and np.array_equal(d, g[f"d{order}"])
    p, v = orc.leapfrog(g["m"], g["pos"], g["vel"], 10, 1e-3, order, 0.0)
    assert np.max(np.abs(p - g[f"pos_lf{order}"])) < 1e-13 * (1 + np.max(np.abs(g[f"pos_lf{order}"])))
    assert np.max(np.abs(v - g[f"vel_lf{order}"])) < 1e-12 * np.max(np.abs(g[f"vel_lf{order}"]))


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_leapfrog_matches_reference_close_encounters(orc, order):
    # 300 jittered bodies with close encounters: chaotic, so the bound is looser
    g = golden("jitter_300.npz")
    p, v = orc.leapfrog(g["m"], g["pos"], g["vel"], 10, 1e-3, order, 0.0)
    assert np.max(np.abs(p - g[f"pos_lf{order}"])) < 1e-9 * (1 + np.max(np.abs(g[f"pos_lf{order}"])))
    assert np.max(np.abs(v - g[f"vel_lf{order}"])) < 1e-7 * np.max(np.abs(g[f"vel_lf{order}"]))


def test_reference_fp32_form_error_level():
    # the reference's only reduced-precision form (compute_forces_avx2_blocked_sp)
    # sits at ~1e-6..1e-5 relative force error; the mixed form is held to this level
    g = golden("jitter_300.npz")
    per = np.linalg.norm(g["force_sp"] - g["force_avx2"], axis=0) / np.linalg.norm(g["force_avx2"], axis=0)
    assert 1e-8 < per.max() < 1e-4


# --- live checks against the compiled reference (skipped when it is absent) ---

def test_live_ring_ic_bitwise(orc, ref):
    for n in (2, 3, 64, 1001):
        a = orc.ic_ring(n, 1.0, 2 * np.pi, 1e5)
        b = ref.init_ring(n, 1.0, 2 * np.pi, 1e5)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_live_plummer_forces(orc, ref):
    m, pos, vel = orc.ic_plummer(1500, 7)
    f = orc.accel(m, pos, 0.0) * m
    for variant in (ref.AVX2_BLOCKED, ref.AVX512_RSQRT_4I, ref.AVX512_RSQRT, ref.AVX2):
        fr = ref.forces(variant, m, pos)
        per = np.linalg.norm(f - fr, axis=0) / np.linalg.norm(fr, axis=0)
        assert per.max() < 1e-11, variant


def test_live_leapfrog_sphere(orc, ref):
    m, pos, vel = orc.ic_sphere(256, 3)
    c, d, lf = orc.yoshida(4)
    p1, v1 = ref.leapfrog(m, pos, vel, 5, 1e-3, c, d)
    p2, v2 = orc.leapfrog(m, pos, vel, 5, 1e-3, 4, 0.0)
    assert np.max(np.abs(p1 - p2)) < 1e-13
    assert np.max(np.abs(v1 - v2)) < 1e-11 * np.max(np.abs(v1))
