"""GPU Euler and Runge-Kutta drivers, and the CLI's run types, against the
reference's fixtures and the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import nbody_amd
from conftest import ROOT, golden

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "n-body-ring-solver_amd", "bin", "nbody")
TWO_PI = 2.0 * np.pi


def close(a, b, rtol):
    return float(np.max(np.abs(a - b))) <= rtol * (1.0 + float(np.max(np.abs(b))))


def run_driver(solver, kind, m, pos, vel, steps, dt, form="fp64", eps=0.0, kernel="auto", tableau=None):
    solver.set_kernel(kernel)
    try:
        solver.set_state(m, pos, vel)
        if kind == "euler":
            solver.step_euler(steps, dt, form, eps)
        else:
            solver.step_rk(steps, dt, form, eps, tableau)
        return solver.get_state()
    finally:
        solver.set_kernel("auto")


@pytest.mark.parametrize("kernel", ["tile", "symmetric"])
@pytest.mark.parametrize("n,tol", [(64, 1e-12), (300, 1e-8)])
@pytest.mark.parametrize("kind", ["euler", "rk"])
def test_reference_driver_fixtures(solver, kind, n, tol, kernel):
    g = golden(f"drivers_{n}.npz")
    tab = (g["rk_a"], g["rk_b"], 4) if kind == "rk" else None
    p, v = run_driver(solver, kind, g["m"], g["pos"], g["vel"], 10, 1e-3, kernel=kernel, tableau=tab)
    key = "euler" if kind == "euler" else "rk4"
    assert close(p, g[f"pos_{key}"], tol)
    assert close(v, g[f"vel_{key}"], 100 * tol)


@pytest.mark.parametrize("n", [1000, 40000])
@pytest.mark.parametrize("kind", ["euler", "rk"])
def test_drivers_vs_oracle(solver, orc, kind, n):
    m, pos, vel, _ = nbody_amd.init_system("plummer", n, 5)
    steps = 4 if n > 10000 else 12
    p, v = run_driver(solver, kind, m, pos, vel, steps, 1.0 / 128, "fp64", 0.01)
    if kind == "euler":
        po, vo = orc.euler(m, pos, vel, steps, 1.0 / 128, 0.01)
    else:
        po, vo = orc.rk(m, pos, vel, steps, 1.0 / 128, eps=0.01)
    assert np.max(np.abs(p - po)) < 1e-12 * np.max(np.abs(po))
    assert np.max(np.abs(v - vo)) < 1e-11 * np.max(np.abs(vo))


def test_one_stage_tableau_is_euler_bitwise(solver):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 777, 8)
    pe, ve = run_driver(solver, "euler", m, pos, vel, 9, 1e-3, "fp64", 0.02)
    pr, vr = run_driver(solver, "rk", m, pos, vel, 9, 1e-3, "fp64", 0.02, tableau=(np.zeros(0), np.array([1.0]), 1))
    assert np.array_equal(pe, pr) and np.array_equal(ve, vr)


def test_rk4_ring_quarter_revolution(solver, orc):
    n = 129
    m, pos, vel, _ = nbody_amd.init_system("ring", n, 0, [1.0, TWO_PI, 1e5])
    p, v = run_driver(solver, "rk", m, pos, vel, 100, 0.25 / 100)
    pp, vp = orc.ring_predict(n, 1.0, TWO_PI, 0.25)
    assert np.max(np.abs(p - pp)) < 1e-7 and np.max(np.abs(v - vp)) < 1e-6


@pytest.mark.parametrize("kind", ["euler", "rk"])
def test_mixed_drivers_track_fp64(solver, kind):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 4096, 3)
    p64, v64 = run_driver(solver, kind, m, pos, vel, 8, 1.0 / 256, "fp64", 0.01)
    pmx, vmx = run_driver(solver, kind, m, pos, vel, 8, 1.0 / 256, "mixed", 0.01)
    assert np.max(np.abs(pmx - p64)) < 1e-8
    assert np.max(np.abs(vmx - v64)) < 1e-5 * np.max(np.abs(v64))


def test_driver_timing_and_errors(solver):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 2048, 2)
    solver.set_state(m, pos, vel)
    solver.timing(True)
    solver.timing_reset()
    solver.step_euler(3, 1e-3, "fp64", 0.01)
    solver.step_rk(2, 1e-3, "fp64", 0.01)
    ms, evals, step_ms = solver.timing_read()
    solver.timing(False)
    assert evals == 3 + 2 * 4 and ms > 0 and step_ms >= ms
    with pytest.raises(nbody_amd.NBodyError):
        solver.step_euler(1, 1e-3, "learned", 0.0)
    with pytest.raises(nbody_amd.NBodyError):
        solver.step_rk(1, 1e-3, "fp64", 0.0, (np.zeros(81), np.ones(9) / 9, 9))
    with pytest.raises(nbody_amd.NBodyError):
        solver.step_euler(1, float("nan"), "fp64", 0.0)
    p0, _ = solver.get_state()
    solver.step_euler(0, 1e-3, "fp64", 0.0)
    assert np.array_equal(solver.get_state()[0], p0)


def test_other_drivers_restart_learned_history(solver):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 2048, 14)
    solver.learned_configure(period=4, threshold=1e-2, exact_form="fp64")
    solver.learned_set_mode(1)
    solver.set_state(m, pos, vel)
    solver.step(24, 1.0 / 256, 2, "learned", 0.01)
    st1 = solver.learned_stats()
    assert st1["predicted_evals"] > 0
    solver.step_euler(1, 1.0 / 256, "fp64", 0.01)   # the state moves outside the learned history
    solver.step(3, 1.0 / 256, 2, "learned", 0.01)    # so these substeps are exact
    st2 = solver.learned_stats()
    assert st2["predicted_evals"] == st1["predicted_evals"]
    assert st2["exact_evals"] - st1["exact_evals"] == 3


# ---------------------------------------------------------------------------
# command line
# ---------------------------------------------------------------------------

def cli(*args, cwd=None):
    print("CLI:", " ".join(args), flush=True)
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120, cwd=cwd)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


@pytest.mark.parametrize("integ,args", [("euler", []), ("lf", ["-i", "lf", "-order", "4"]),
                                        ("rk", ["-i", "rk", "-order", "4"])])
def test_cli_integrators(orc, tmp_path, integ, args):
    snap = str(tmp_path / "s.dat")
    out = cli("-t", "omp", "-N", "64", "-s", "10", "-T", "0.01", "-snapshot", snap, *args)
    m, pos, vel, _ = orc.ic_ring(64, 1.0, TWO_PI, 1e5)
    if integ == "euler":
        assert "Reduced Euler blocked solver took" in out
        po, vo = orc.euler(m, pos, vel, 10, 1e-3)
    elif integ == "lf":
        assert "Reduced Symplectic Solver took" in out and "c[3] = " in out
        po, vo = orc.leapfrog(m, pos, vel, 10, 1e-3, 4)
    else:
        assert "Reduced Runge-Kutta Solver took" in out
        po, vo = orc.rk(m, pos, vel, 10, 1e-3)
    assert nbody_amd.snapshot_frames(snap, 64) == 2
    _, _, p, v, _ = nbody_amd.snapshot_read(snap, 64, 1)
    assert np.max(np.abs(p - po)) < 1e-12 and np.max(np.abs(v - vo)) < 1e-10
    assert "Performance:" in out and "Average position deviation from reference" in out


def test_cli_default_is_euler_order_fallbacks(tmp_path):
    out = cli("-N", "32", "-s", "2", "-i", "rk", "-order", "3")
    assert "Order 3 not supported! Defaulting to 4th order" in out
    out = cli("-N", "32", "-s", "2", "-i", "lf")
    assert "Order 1 not supported! Defaulting to 2nd order" in out
    out = cli("-N", "32", "-s", "2", "-S", "full")
    assert "Reduced Euler blocked solver took" in out


def test_cli_accuracy_experiment(tmp_path):
    csv = str(tmp_path / "acc.csv")
    for prec in ("dp", "sp"):
        cli("-t", "omp-accuracy", "-start", "256", "-step", "2", "-n", "2", "-ratio-start", "1e3",
            "-ratio-stop", "1e6", "-ratio-steps", "1", "-p", prec, "-o", csv)
    lines = open(csv).read().strip().splitlines()
    assert lines[0] == "N, precision, ratio, average_force_length_error[%], average_force_angle_error[%]"
    rows = [[x.strip() for x in ln.split(",")] for ln in lines[1:]]
    assert len(rows) == 8
    assert [r[0] for r in rows[:4]] == ["256", "512", "256", "512"]
    assert [float(r[2]) for r in rows[:4]] == [1e3, 1e3, 1e6, 1e6]
    for r in rows:
        length, angle = float(r[3]), float(r[4])
        if r[1] == "dp":
            assert length < 1e-10 and angle < 1e-10        # percent; FP64 sums
        else:
            assert 1e-10 < length < 1e-2 and angle < 1e-2   # the all-FP32 form


def test_cli_ratio_experiment_vs_oracle(orc, tmp_path):
    """Permutation-invariant quantities of the rigid-rotation experiment
    against the oracle run on the unpermuted ring."""
    csv = str(tmp_path / "ratio.csv")
    n, s, rev = 64, 40, 200
    out = cli("-t", "omp-ratio", "-start", str(n), "-n", "1", "-s", str(s), "-rev-steps", str(rev),
              "-ratio-start", "1e6", "-ratio-stop", "1e6", "-ratio-steps", "0", "-max-error", "0.5", "-o", csv)
    lines = open(csv).read().strip().splitlines()
    assert lines[0].startswith("N, precision, s, s_collapse, ratio, max_error, last_E_dev")
    row = [x.strip() for x in lines[1].split(",")]
    assert row[:4] == [str(n), "dp", str(s), str(s)]
    got = re.search(r"last E_dev = (\S+), last max pos dev = (\S+), last max vel dev = (\S+)", out)
    e_dev, pmax, vmax = (float(x) for x in got.groups())
    m, pos, vel, _ = orc.ic_ring(n, 1.0, TWO_PI, 1e6)
    T0, U0 = orc.energy(m, pos, vel, 0.0)
    p, v = orc.leapfrog(m, pos, vel, s, 1.0 / rev, 2)
    T, U = orc.energy(m, p, v, 0.0)
    pp, vp = nbody_amd.ring_predict(n, 1.0, TWO_PI, s, rev)
    want_e = abs((T0 + U0 - T - U) / (T0 + U0))
    assert abs(e_dev - want_e) <= 1e-6 * want_e + 1e-15
    assert abs(pmax - orc.deviation(p, pp)[1]) <= 1e-6 * pmax
    assert abs(vmax - orc.deviation(v, vp)[1]) <= 1e-6 * vmax


def test_cli_csv_row_and_verbose(tmp_path):
    csv = str(tmp_path / "run.csv")
    out = cli("-t", "omp", "-N", "16", "-s", "3", "-T", "0.003", "-i", "lf", "-order", "2", "-v", "-o", csv)
    assert "calc_position[15]=" in out and "ref_velocity[0] =" in out and "velocity_diff[0] =" in out
    lines = open(csv).read().strip().splitlines()
    assert lines[0].startswith("N, s, T, dt, integrator, order, eps")
    row = [x.strip() for x in lines[1].split(",")]
    assert row[0] == "16" and row[1] == "3" and row[4] == "lf" and row[5] == "2" and row[9] == "fp64"
    cli("-t", "omp", "-N", "16", "-s", "3", "-T", "0.003", "-i", "rk", "-order", "4", "-p", "sp", "-o", csv)
    row = [x.strip() for x in open(csv).read().strip().splitlines()[2].split(",")]
    assert row[4] == "rk" and row[5] == "4" and row[9] == "fp32"   # -p sp: the all-FP32 form


def test_cli_learned_form():
    out = cli("-t", "omp", "-N", "4096", "-ic", "plummer", "-eps", "0.01", "-s", "40", "-T", str(40 / 256),
              "-form", "learned", "-train-steps", "24", "-train-iters", "50")
    assert "using -i lf" in out and "Training the extrapolator" in out and "final training loss" in out
    e = float(re.search(r"\|dE/E\| = (\S+)", out).group(1))
    assert e < 1e-4
