"""Euler and Runge-Kutta drivers and the accuracy-experiment metrics (CPU only).

Pins the oracle's restatements of the reference's Euler driver
(reduced_solver_omp, CPU/omp_solver.cpp:16-137), its Runge-Kutta driver
(reduced_solver_omp_rk, :458-606), and the metrics of its accuracy
experiments (calculate_force_error, calculate_deviation, compute_inertia,
predict_stable_orbiting_particles) on the golden fixtures written from the
reference's compiled routines, on closed forms (convergence orders, rotated
and scaled force fields), and live against the reference where oracle/_ref is
built.  The library's host-side metric functions (nb_force_error,
nb_deviation, nb_inertia, nb_ring_predict, nb_runge_kutta) need no GPU and
are checked here bit for bit.
"""
import numpy as np
import pytest

import nbody_amd
from conftest import golden
from test_oracle_kats import two_body

TWO_PI = 2.0 * np.pi


def close(a, b, rtol):
    return float(np.max(np.abs(a - b))) <= rtol * (1.0 + float(np.max(np.abs(b))))


# ---------------------------------------------------------------------------
# drivers against the reference's own trajectories
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n,tol", [(64, 1e-13), (300, 1e-9)])
def test_euler_matches_reference_driver(orc, n, tol):
    g = golden(f"drivers_{n}.npz")
    p, v = orc.euler(g["m"], g["pos"], g["vel"], 10, 1e-3)
    assert close(p, g["pos_euler"], tol)
    assert close(v, g["vel_euler"], 100 * tol)


@pytest.mark.parametrize("n,tol", [(64, 1e-13), (300, 1e-9)])
def test_rk4_matches_reference_driver(orc, n, tol):
    g = golden(f"drivers_{n}.npz")
    p, v = orc.rk(g["m"], g["pos"], g["vel"], 10, 1e-3, g["rk_a"], g["rk_b"])
    assert close(p, g["pos_rk4"], tol)
    assert close(v, g["vel_rk4"], 100 * tol)


def test_runge_kutta_tableau():
    g = golden("drivers_64.npz")
    for order in (4, 1, 2, 7):   # every order falls back to RK4, as the reference's switch does
        a, b, s = nbody_amd.runge_kutta(order)
        assert s == 4
        assert np.array_equal(a, g["rk_a"]) and np.array_equal(b, g["rk_b"])
    assert abs(b.sum() - 1.0) < 1e-15


# ---------------------------------------------------------------------------
# closed forms: convergence orders on the circular binary
# ---------------------------------------------------------------------------

def period_error(step, m, pos, vel, P, steps):
    p, _ = step(m, pos, vel, steps, P / steps)
    return float(np.max(np.abs(p - pos)))   # one full period returns to the start


EPS = 0.2   # the binary's softening (two_body builds the circular orbit for it)


def test_euler_first_order(orc):
    m, pos, vel, P = two_body(EPS)
    euler = lambda *x: orc.euler(*x, eps=EPS)   # noqa: E731
    e1 = period_error(euler, m, pos, vel, P, 4000)
    e2 = period_error(euler, m, pos, vel, P, 8000)
    assert abs(np.log2(e1 / e2) - 1.0) < 0.15, (e1, e2)


def test_rk4_fourth_order(orc):
    m, pos, vel, P = two_body(EPS)
    rk = lambda *x: orc.rk(*x, eps=EPS)   # noqa: E731
    e1 = period_error(rk, m, pos, vel, P, 160)
    e2 = period_error(rk, m, pos, vel, P, 320)
    assert abs(np.log2(e1 / e2) - 4.0) < 0.25, (e1, e2)


def test_rk_general_tableaus(orc):
    m, pos, vel, P = two_body(EPS)
    # one stage, b = [1]: the Euler step exactly
    p1, v1 = orc.rk(m, pos, vel, 7, P / 50, np.zeros(0), np.array([1.0]), eps=EPS)
    pe, ve = orc.euler(m, pos, vel, 7, P / 50, eps=EPS)
    assert np.array_equal(p1, pe) and np.array_equal(v1, ve)
    # explicit midpoint (a = [1/2], b = [0, 1]): second order
    mid = (np.array([0.5]), np.array([0.0, 1.0]))
    errs = [period_error(lambda *x: orc.rk(*x, *mid, eps=EPS), m, pos, vel, P, s) for s in (200, 400)]
    assert abs(np.log2(errs[0] / errs[1]) - 2.0) < 0.3, errs


# ---------------------------------------------------------------------------
# planted mistakes in numpy restatements: the fixture checks reject them
# ---------------------------------------------------------------------------

def numpy_euler(orc, m, pos, vel, steps, dt, mistake=None):
    x, v = pos.copy(), vel.copy()
    for _ in range(steps):
        a = orc.accel(m, np.ascontiguousarray(x), 0.0)
        if mistake == "kick_first":
            v = v + a * dt
            x = x + v * dt
        else:
            x = x + v * dt
            v = v + a * dt
    return x, v


def numpy_rk4(orc, m, pos, vel, steps, dt, mistake=None):
    A = [[0.5], [0.0, 0.5], [0.0, 0.0, 1.0]]
    b = [1 / 6, 1 / 3, 1 / 3, 1 / 6]
    x, v = pos.copy(), vel.copy()
    for _ in range(steps):
        kv, ka = [v], []
        xs = x
        for k in range(4):
            ka.append(orc.accel(m, np.ascontiguousarray(xs), 0.0))
            if k == 3:
                break
            base = xs if mistake == "stage_from_stage" else x
            xs = base + dt * sum(w * kv[j] for j, w in enumerate(A[k]))
            nv = v + dt * sum(w * ka[j] for j, w in enumerate(A[k]))
            kv.append(v if mistake == "stale_stage_velocity" else nv)
        x = x + dt * sum(w * kv[j] for j, w in enumerate(b))
        v = v + dt * sum(w * ka[j] for j, w in enumerate(b))
    return x, v


@pytest.mark.parametrize("mistake", [None, "kick_first"])
def test_planted_euler_mistake_is_caught(orc, mistake):
    g = golden("drivers_64.npz")
    p, v = numpy_euler(orc, g["m"], g["pos"], g["vel"], 10, 1e-3, mistake)
    ok = close(p, g["pos_euler"], 1e-12) and close(v, g["vel_euler"], 1e-11)
    assert ok == (mistake is None)


@pytest.mark.parametrize("mistake", [None, "stage_from_stage", "stale_stage_velocity"])
def test_planted_rk_mistake_is_caught(orc, mistake):
    g = golden("drivers_64.npz")
    p, v = numpy_rk4(orc, g["m"], g["pos"], g["vel"], 10, 1e-3, mistake)
    ok = close(p, g["pos_rk4"], 1e-12) and close(v, g["vel_rk4"], 1e-11)
    assert ok == (mistake is None)


# ---------------------------------------------------------------------------
# accuracy-experiment metrics
# ---------------------------------------------------------------------------

def test_metrics_match_reference_fixture(orc):
    g = golden("accuracy_257.npz")
    for fn in (orc.force_error, nbody_amd.force_error):
        assert np.array_equal(fn(g["force_avx2"], g["force_ref"]), g["force_error"])
    for fn in (orc.deviation, nbody_amd.deviation):
        assert np.array_equal(fn(g["pos_jit"], g["pos"]), g["deviation"])
    assert orc.inertia(g["m"], g["pos_jit"]) == g["inertia"]
    assert nbody_amd.inertia(g["m"], g["pos_jit"]) == g["inertia"]
    p, v = nbody_amd.ring_predict(257, 1.0, TWO_PI, 37, 1000)
    assert np.array_equal(p, g["pos_pred"]) and np.array_equal(v, g["vel_pred"])


def test_force_error_closed_forms(orc):
    # reference forces of unit length in the xy plane at angles phi_i (none on an axis)
    n = 101
    phi = np.linspace(0.1, 6.0, n)
    ref = np.ascontiguousarray(np.stack([np.cos(phi), np.sin(phi), np.zeros(n)]))
    th, s = 1e-3, 1.0 + 2e-4
    c, sn = np.cos(th), np.sin(th)
    rot = np.ascontiguousarray(s * np.stack([c * ref[0] - sn * ref[1], sn * ref[0] + c * ref[1], ref[2]]))
    for fn in (orc.force_error, nbody_amd.force_error):
        e = fn(rot, ref)
        chord = np.hypot(s * c - 1.0, s * sn)        # |s R(th) u - u| for a unit u
        assert abs(e[0] - chord) < 1e-15 and abs(e[1] - chord) < 1e-15
        assert abs(e[2] - th / TWO_PI) < 1e-15 and abs(e[4] - th / TWO_PI) < 1e-15
        L0 = float(np.float32(np.sqrt(np.float32(ref[0, 0]) ** 2 + np.float32(ref[1, 0]) ** 2)))
        assert abs(e[3] - (s - 1.0) / L0) < 1e-12 and abs(e[5] - (s - 1.0) / L0) < 1e-12
    # z is ignored by the angle (what is left is the rounding of the contracted
    # determinant, as in the reference); axis-aligned reference forces are left out
    ref2 = ref.copy()
    ref2[1, 0] = 0.0
    tilt = ref2.copy()
    tilt[2] += 0.3
    e = nbody_amd.force_error(tilt, ref2)
    assert e[2] < 1e-16 and e[4] < 1e-16
    assert np.array_equal(e, orc.force_error(tilt, ref2))


def test_deviation_and_inertia_closed_forms(orc):
    rng = np.random.default_rng(5)
    ref = rng.standard_normal((3, 50))
    ref[:, 7] = 0.0                                   # |ref| = 0: left out of the maximum
    calc = ref * 1.5
    calc[:, 7] = 1.0
    d = nbody_amd.deviation(calc, ref)
    assert np.array_equal(d, orc.deviation(calc, ref))
    lens = np.linalg.norm(ref, axis=0)
    assert abs(d[0] - (0.5 * lens.sum() + np.sqrt(3.0)) / 50) < 1e-14
    assert abs(d[1] - 0.5) < 1e-15
    m = rng.uniform(0.5, 2.0, 50)
    assert abs(nbody_amd.inertia(m, ref) - float((m * lens ** 2).sum())) < 1e-12


def test_ring_predict_quarter_turn(orc):
    n = 17
    p0, v0 = nbody_amd.ring_predict(n, 2.0, TWO_PI, 0, 4)
    m, pos, vel, _ = orc.ic_ring(n, 2.0, TWO_PI, 1e4)
    assert np.max(np.abs(p0 - pos)) < 1e-15 and np.max(np.abs(v0 - vel)) < 1e-14
    p1, v1 = nbody_amd.ring_predict(n, 2.0, TWO_PI, 1, 4)   # a quarter turn: (x, y) -> (-y, x)
    assert np.max(np.abs(p1[0] + pos[1])) < 1e-14 and np.max(np.abs(p1[1] - pos[0])) < 1e-14
    assert np.all(p1[:, -1] == 0.0) and np.all(v1[:, -1] == 0.0)
    po, _ = orc.ring_predict(n, 2.0, TWO_PI, 0.25)
    assert np.max(np.abs(po - p1)) < 1e-15


def test_metric_argument_errors():
    z = np.zeros((3, 4))
    with pytest.raises(nbody_amd.NBodyError):
        nbody_amd.force_error(np.zeros((3, 0)), np.zeros((3, 0)))
    with pytest.raises(nbody_amd.NBodyError):
        nbody_amd.deviation(np.zeros((3, 0)), np.zeros((3, 0)))
    with pytest.raises(nbody_amd.NBodyError):
        nbody_amd.ring_predict(1, 1.0, 1.0, 0, 10)
    with pytest.raises(nbody_amd.NBodyError):
        nbody_amd.ring_predict(8, 1.0, 1.0, 0, 0)
    assert nbody_amd.inertia(np.zeros(0), np.zeros((3, 0))) == 0.0
    assert nbody_amd.deviation(z, z)[1] == 0.0


# ---------------------------------------------------------------------------
# live against the compiled reference (skipped when it is absent)
# ---------------------------------------------------------------------------

def test_live_drivers(orc, ref):
    m, pos, vel = orc.ic_plummer(48, 3)
    pr, vr = ref.euler(m, pos, vel, 5, 2e-3)
    po, vo = orc.euler(m, pos, vel, 5, 2e-3)
    assert close(po, pr, 1e-13) and close(vo, vr, 1e-12)
    a, b, _ = nbody_amd.runge_kutta(4)
    pr, vr = ref.rk(m, pos, vel, 5, 2e-3, a, b)
    po, vo = orc.rk(m, pos, vel, 5, 2e-3, a, b)
    assert close(po, pr, 1e-13) and close(vo, vr, 1e-12)


def test_live_metrics_bitwise(orc, ref):
    rng = np.random.default_rng(11)
    for n in (1, 2, 33, 500):
        r = rng.standard_normal((3, n))
        c = r + 1e-5 * rng.standard_normal((3, n))
        c[:, 0] = r[:, 0] * 2.0
        if n > 2:
            r[0, 1] = 0.0                           # an axis-aligned reference force
        assert np.array_equal(nbody_amd.force_error(c, r), ref.force_error(c, r))
        assert np.array_equal(nbody_amd.deviation(c, r), ref.deviation(c, r))
        mm = rng.uniform(0.1, 1.0, n)
        assert nbody_amd.inertia(mm, c) == ref.inertia(mm, c)
    for n, step, rev in ((2, 1, 3), (64, 999, 1000), (1001, 12345, 777)):
        p, v = nbody_amd.ring_predict(n, 1.0, TWO_PI, step, rev)
        pr, vr = ref.ring_predict(n, 1.0, TWO_PI, step, rev)
        assert np.array_equal(p, pr) and np.array_equal(v, vr)
