"""The whole-run entry point in the FPGA host's record form (nb_run_records,
the counterpart of lrbd_solver_no_sync, FPGA/host/solver.cpp:1112): AoS
records of 4 doubles advanced in place, against the oracle's leapfrog and
force; the scalar factor K and the pair weights; argument errors leave the
resident state alone."""
import numpy as np
import pytest

import nbody_amd
from conftest import rel_err

pytestmark = pytest.mark.gpu

SENTINEL = -7.25   # the records' 4th word: never read, never written


def records(a):
    r = np.full((a.shape[1], 4), SENTINEL)
    r[:, :3] = a.T
    return r


@pytest.mark.parametrize("order", [2, 4])
def test_run_records_vs_oracle(solver, orc, order):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 512, 11)
    steps, dt, eps = 10, 1.0 / 128, 0.01
    p4, v4, f4 = records(pos), records(vel), np.full((512, 4), SENTINEL)
    secs = solver.run_records(m, p4, v4, steps, dt, order, "fp64", eps, force4=f4)
    assert secs > 0.0
    po, vo = orc.leapfrog(m, pos, vel, steps, dt, order, eps)
    assert np.max(np.abs(p4[:, :3].T - po)) < 1e-12 * np.max(np.abs(po))
    assert np.max(np.abs(v4[:, :3].T - vo)) < 1e-11 * np.max(np.abs(vo))
    for r in (p4, v4, f4):
        assert np.all(r[:, 3] == SENTINEL)
    if order == 2:
        # the last force evaluation sees the positions after the last step's
        # first half drift (c = 1/2)
        pp, vp = orc.leapfrog(m, pos, vel, steps - 1, dt, order, eps)
        x_mid = pp + 0.5 * dt * vp
        f_ref = m * orc.accel(m, x_mid, eps)
        assert rel_err(f4[:, :3].T, f_ref) < 1e-9


def test_run_records_unequal_masses_odd_n(solver, orc):
    # masses apart (the kernels' general path) and an odd N; coeff = mass
    rng = np.random.default_rng(4)
    n = 777
    _, pos, vel, _ = nbody_amd.init_system("plummer", n, 13)
    m = rng.uniform(0.5, 1.5, n) / n
    p4, v4, f4 = records(pos), records(vel), np.zeros((n, 4))
    solver.run_records(m, p4, v4, 6, 1.0 / 128, 2, "fp64", 0.01, coeff=m.copy(), force4=f4)
    po, vo = orc.leapfrog(m, pos, vel, 6, 1.0 / 128, 2, 0.01)
    assert np.max(np.abs(p4[:, :3].T - po)) < 1e-12 * np.max(np.abs(po))
    assert np.max(np.abs(v4[:, :3].T - vo)) < 1e-11 * np.max(np.abs(vo))
    pp, vp = orc.leapfrog(m, pos, vel, 5, 1.0 / 128, 2, 0.01)
    f_ref = m * orc.accel(m, pp + 0.5 / 128 * vp, 0.01)
    assert rel_err(f4[:, :3].T, f_ref) < 1e-9


def test_run_records_equals_the_resident_state_run(solver):
    # the same steps through nb_set_state / nb_step / nb_get_state: bitwise
    m, pos, vel, _ = nbody_amd.init_system("sphere", 3000, 2)
    p4, v4 = records(pos), records(vel)
    solver.run_records(m, p4, v4, 5, 1e-3, 4, "mixed", 1e-3)
    solver.set_state(m, pos, vel)
    solver.step(5, 1e-3, 4, "mixed", 1e-3)
    p, v = solver.get_state()
    assert np.array_equal(p4[:, :3].T, p) and np.array_equal(v4[:, :3].T, v)


def test_run_records_scalar_factor(solver):
    # K = 2 with coeff = mass = m is the same dynamics as K = 1 with
    # coeff = mass = 2 m; the force (K coeff_i coeff_j ...) doubles
    m, pos, vel, _ = nbody_amd.init_system("plummer", 2000, 3)
    pa, va, fa = records(pos), records(vel), np.zeros((2000, 4))
    solver.run_records(m, pa, va, 6, 1.0 / 256, 2, "fp64", 0.005, K=2.0, force4=fa)
    pb, vb, fb = records(pos), records(vel), np.zeros((2000, 4))
    solver.run_records(2 * m, pb, vb, 6, 1.0 / 256, 2, "fp64", 0.005, K=1.0, force4=fb)
    assert np.max(np.abs(pa - pb)) < 1e-13 * np.max(np.abs(pb))
    assert np.max(np.abs(va - vb)) < 1e-12 * np.max(np.abs(vb))
    assert rel_err(2 * fa[:, :3].T, fb[:, :3].T) < 1e-12
    # pair weights apart from the masses: the weight enters every pair twice
    # (K coeff_i coeff_j), so coeff = 3 m with K = 1/9 is the dynamics of
    # coeff = mass = m (the kick divides by the mass)
    pc, vc = records(pos), records(vel)
    solver.run_records(m, pc, vc, 6, 1.0 / 256, 2, "fp64", 0.005, K=1.0 / 9.0, coeff=3 * m)
    pd, vd = records(pos), records(vel)
    solver.run_records(m, pd, vd, 6, 1.0 / 256, 2, "fp64", 0.005)
    assert np.max(np.abs(pc - pd)) < 1e-13 * np.max(np.abs(pd))
    assert np.max(np.abs(vc - vd)) < 1e-12 * np.max(np.abs(vd))


def test_run_records_fp32_form(solver):
    # the all-FP32 form stands in for lrbd_solver_sp
    m, pos, vel, _ = nbody_amd.init_system("plummer", 4096, 5)
    p64, v64 = records(pos), records(vel)
    solver.run_records(m, p64, v64, 4, 1.0 / 256, 2, "fp64", 0.005)
    p32, v32 = records(pos), records(vel)
    solver.run_records(m, p32, v32, 4, 1.0 / 256, 2, "fp32", 0.005)
    assert 0.0 < np.max(np.abs(v32 - v64)) < 1e-4 * np.max(np.abs(v64))


def test_run_records_errors_leave_the_state(solver):
    m, pos, vel, _ = nbody_amd.init_system("sphere", 256, 9)
    solver.set_state(m, pos, vel)
    p_before, v_before = solver.get_state()
    p4, v4 = records(pos), records(vel)
    coeff = m.copy()
    coeff[17] *= 1.5   # K coeff_i / mass_i no longer one value
    with pytest.raises(nbody_amd.NBodyError):
        solver.run_records(m, p4, v4, 3, 1e-3, 2, "fp64", 1e-3, coeff=coeff)
    bad = m.copy()
    bad[3] = 0.0       # a zero mass has no kick
    with pytest.raises(nbody_amd.NBodyError):
        solver.run_records(bad, p4, v4, 3, 1e-3, 2, "fp64", 1e-3)
    with pytest.raises(nbody_amd.NBodyError):
        solver.run_records(m, p4, v4, 3, 1e-3, 2, "fp64", -1.0)
    with pytest.raises(nbody_amd.NBodyError):   # K coeff / mass overflows
        solver.run_records(m, p4, v4, 3, 1e-3, 2, "fp64", 1e-3, K=1e308, coeff=1e10 * m)
    assert np.all(p4 == records(pos)) and np.all(v4 == records(vel))
    p, v = solver.get_state()
    assert np.array_equal(p, p_before) and np.array_equal(v, v_before)
    # zero steps change nothing
    assert solver.run_records(m, p4, v4, 0, 1e-3, 2, "fp64", 1e-3) == 0.0
    assert np.all(p4 == records(pos))
