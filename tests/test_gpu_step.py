"""GPU leapfrog and energy parity against the oracle and the reference's
leapfrog fixtures; integrator invariants on the GPU."""
import numpy as np
import pytest

import nbody_amd
from conftest import golden

pytestmark = pytest.mark.gpu


def run(solver, m, pos, vel, steps, dt, order, form="fp64", eps=0.0):
    solver.set_state(m, pos, vel)
    solver.step(steps, dt, order, form, eps)
    return solver.get_state()


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_reference_leapfrog_fixture(solver, order):
    g = golden("jitter_64.npz")
    p, v = run(solver, g["m"], g["pos"], g["vel"], 10, 1e-3, order)
    assert np.max(np.abs(p - g[f"pos_lf{order}"])) < 1e-12 * (1 + np.max(np.abs(g[f"pos_lf{order}"])))
    assert np.max(np.abs(v - g[f"vel_lf{order}"])) < 1e-11 * np.max(np.abs(g[f"vel_lf{order}"]))


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_leapfrog_vs_oracle(solver, orc, order):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 512, 4)
    p, v = run(solver, m, pos, vel, 10, 1.0 / 128, order, "fp64", 0.01)
    po, vo = orc.leapfrog(m, pos, vel, 10, 1.0 / 128, order, 0.01)
    assert np.max(np.abs(p - po)) < 1e-12 * np.max(np.abs(po))
    assert np.max(np.abs(v - vo)) < 1e-11 * np.max(np.abs(vo))


def test_cold_collapse_config(solver, orc):
    # N = 2048 uniform sphere, eps = 1e-3, dt = 1e-3, 100 steps (-T 0.1 -s 100)
    m, pos, vel, _ = nbody_amd.init_system("sphere", 2048, 42)
    p, v = run(solver, m, pos, vel, 100, 1e-3, 2, "fp64", 1e-3)
    po, vo = orc.leapfrog(m, pos, vel, 100, 1e-3, 2, 1e-3)
    assert np.max(np.abs(p - po)) < 1e-10
    assert np.max(np.abs(v - vo)) < 1e-8 * np.max(np.abs(vo))


def test_ring_quarter_revolution(solver, orc):
    n = 129
    m, pos, vel, _ = nbody_amd.init_system("ring", n, 0, [1.0, 2 * np.pi, 1e5])
    p, v = run(solver, m, pos, vel, 400, 0.25 / 400, 4)
    pp, vp = orc.ring_predict(n, 1.0, 2 * np.pi, 0.25)
    assert np.max(np.abs(p - pp)) < 1e-7 and np.max(np.abs(v - vp)) < 1e-6


def test_time_reversibility(solver):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 1024, 21)
    p1, v1 = run(solver, m, pos, vel, 40, 1e-3, 4, "fp64", 0.05)
    p2, v2 = run(solver, m, p1, np.ascontiguousarray(-v1), 40, 1e-3, 4, "fp64", 0.05)
    assert np.max(np.abs(p2 - pos)) < 1e-12
    assert np.max(np.abs(-v2 - vel)) < 1e-11


def test_mixed_trajectory_tracks_fp64(solver):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 4096, 8)
    p64, v64 = run(solver, m, pos, vel, 20, 1.0 / 128, 2, "fp64", 0.01)
    pm, vm = run(solver, m, pos, vel, 20, 1.0 / 128, 2, "mixed", 0.01)
    assert np.max(np.abs(pm - p64)) < 1e-7


@pytest.mark.parametrize("eps", [0.0, 0.01])
def test_energy_vs_oracle(solver, orc, eps):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 3000, 6)
    m = m * np.linspace(0.5, 1.5, 3000)
    solver.set_state(m, pos, vel)
    T, U = solver.energy(eps)
    To, Uo = orc.energy(m, pos, vel, eps)
    assert abs(T - To) <= 1e-15 * abs(To)
    assert abs(U - Uo) <= 1e-14 * abs(Uo)
    assert solver.energy(eps) == (T, U)   # bitwise reproducible


@pytest.mark.parametrize("n", [1, 2, 255, 257, 700])
@pytest.mark.parametrize("eps", [0.0, 0.01])
def test_energy_ragged_and_coincident(solver, orc, n, eps):
    # tile counts 1..3 with partial last tiles, varied masses and (n >= 4)
    # coincident particles: skipped when eps = 0, m_i m_j / eps otherwise
    m, pos, vel, _ = nbody_amd.init_system("plummer", n, 40 + n)
    m = m * np.linspace(0.5, 1.5, n)
    if n >= 4:
        pos[:, 3] = pos[:, 1]
        pos[:, n - 1] = pos[:, 0]
    solver.set_state(m, pos, vel)
    T, U = solver.energy(eps)
    To, Uo = orc.energy(m, pos, vel, eps)
    assert abs(T - To) <= 1e-15 * abs(To)
    if n == 1:
        assert U == 0.0 and Uo == 0.0
    else:
        assert abs(U - Uo) <= 1e-14 * abs(Uo)


def test_energy_fixture(solver):
    g = golden("jitter_300.npz")
    solver.set_state(g["m"], g["pos"], g["vel"])
    T, U = solver.energy(0.0)
    assert abs(T - g["T"]) <= 1e-14 * abs(g["T"])
    assert abs(U - g["U_scalar"]) <= 1e-14 * abs(g["U_scalar"])


def test_energy_conservation_short(solver):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 8192, 2)
    solver.set_state(m, pos, vel)
    e0 = sum(solver.energy(0.01))
    solver.step(64, 1.0 / 128, 2, "fp64", 0.01)
    e1 = sum(solver.energy(0.01))
    assert abs((e1 - e0) / e0) < 1e-4


def test_step_timing_counters(solver):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 4096, 2)
    solver.set_state(m, pos, vel)
    solver.timing(True)
    solver.timing_reset()
    solver.step(7, 1e-3, 4, "fp64", 0.01)     # order 4: 3 force evaluations per step
    ms, evals, step_ms = solver.timing_read()
    solver.timing(False)
    assert evals == 21 and ms > 0 and step_ms >= ms


@pytest.mark.parametrize("kernel", ["tile", "symmetric"])
def test_leapfrog_both_kernels(solver, orc, kernel):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 3000, 14)
    solver.set_kernel(kernel)
    try:
        p, v = run(solver, m, pos, vel, 8, 1.0 / 128, 4, "fp64", 0.01)
    finally:
        solver.set_kernel("auto")
    po, vo = orc.leapfrog(m, pos, vel, 8, 1.0 / 128, 4, 0.01)
    assert np.max(np.abs(p - po)) < 1e-12 * np.max(np.abs(po))
    assert np.max(np.abs(v - vo)) < 1e-11 * np.max(np.abs(vo))


def momenta(m, pos, vel):
    """Total linear and angular momentum and the sums of their magnitudes
    (the scales the conservation bars are relative to)."""
    P = (m * vel).sum(axis=1)
    Lp = m * np.cross(pos.T, vel.T).T
    return P, Lp.sum(axis=1), (m * np.abs(vel)).sum(), np.abs(Lp).sum()


@pytest.mark.parametrize("n,steps,order", [(32768, 100, 2), (262144, 8, 4)])
@pytest.mark.parametrize("form,tol", [("fp64", 1e-14), ("mixed", 1e-9)])
def test_momentum_and_angular_momentum_conserved(solver, capsys, n, steps, order, form, tol):
    """Pairwise central forces with equal and opposite terms conserve the
    total linear and angular momentum; so does every drift and kick of the
    leapfrog.  Over a run of the symmetric kernels (the item kernel at
    N = 32768, the grouped one at N = 262144) with masses apart, both change
    by no more than rounding: FP64 ~1e-17 of the scales, the mixed form
    (action and reaction terms rounded to FP32 apart) ~1e-11 (both measured
    on one MI355X).  A wrong sign or index in a reaction moves them at O(1);
    a stray acceleration of 1e-3 for one step at dt = 1/128, by ~1e-5."""
    _, pos, vel, _ = nbody_amd.init_system("plummer", n, 31)
    m = np.random.default_rng(n).uniform(0.5, 1.5, n) / n
    P0, L0, sp, sl = momenta(m, pos, vel)
    p, v = run(solver, m, pos, vel, steps, 1.0 / 128, order, form, 0.01)
    P1, L1, _, _ = momenta(m, p, v)
    dp = np.max(np.abs(P1 - P0)) / sp
    dl = np.max(np.abs(L1 - L0)) / sl
    with capsys.disabled():
        print(f"\n  N={n} {form} order {order}, {steps} steps: |dP|/sum|mv| {dp:.2e}, |dL|/sum|m x v| {dl:.2e}")
    assert dp < tol and dl < tol
