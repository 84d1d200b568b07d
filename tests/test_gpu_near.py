"""The near-field split of the learned form on the GPU: the sorted-grid
near-field kernel against the oracle's brute-force sum, and learned stepping
with the split against a numpy restatement of its schedule."""
import numpy as np
import pytest

import nbody_amd

from test_gpu_learned import restated_learned

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,n,rc,eps", [("plummer", 3000, 0.1, 0.01), ("plummer", 3000, 0.4, 0.0),
                                           ("sphere", 2049, 0.25, 1e-3), ("plummer", 700, 100.0, 0.01),
                                           ("plummer", 20000, 0.05, 0.01)])
def test_near_accel_vs_oracle(solver, orc, kind, n, rc, eps):
    # (20000, rc = 0.05) forms ~4100 chunks, more than the resident grid holds
    # (4 blocks per CU on 256 CUs), so blocks take several chunks each
    m, pos, vel, _ = nbody_amd.init_system(kind, n, 3)
    solver.set_state(m, pos, vel)
    a = solver.near_accel(rc, eps)
    ao = orc.near_accel(m, pos, rc, eps)
    scale = np.max(np.linalg.norm(ao, axis=0))
    assert scale > 0
    assert np.max(np.linalg.norm(a - ao, axis=0)) < 1e-12 * scale
    if rc > 80:   # every pair inside rc/2 (radii stay below 20): the whole softened sum
        full = orc.accel(m, pos, eps)
        assert np.max(np.abs(a - full)) < 1e-12 * np.max(np.abs(full))


@pytest.mark.parametrize("rc,n", [(0.1, 4000), (0.3, 4000), (0.1, 40000)])
def test_near_accel_mixed_vs_oracle(solver, orc, rc, n):
    """FP32 pair arithmetic on FP64 differences: within 1e-5 of the largest
    near-field acceleration (FP32 rounding of ~1e3 terms).  N = 40000 at
    rc = 0.1 forms ~3300 chunks (several per resident block)."""
    m, pos, vel, _ = nbody_amd.init_system("plummer", n, 8)
    solver.set_state(m, pos, vel)
    a = solver.near_accel(rc, 0.005, form="mixed")
    ao = orc.near_accel(m, pos, rc, 0.005)
    scale = np.max(np.linalg.norm(ao, axis=0))
    err = np.max(np.linalg.norm(a - ao, axis=0))
    assert err < 1e-5 * scale, err / scale


def test_near_accel_clustered_and_isolated(solver, orc):
    """A dense clump sharing a few cells, plus isolated particles with no
    neighbour inside rc (their near field is exactly zero)."""
    rng = np.random.default_rng(5)
    clump = rng.normal(0.0, 0.01, (3, 500))
    far = np.array([[5.0, -7.0, 3.0], [5.0, 2.0, -9.0], [5.0, 1.0, 4.0]])
    pos = np.ascontiguousarray(np.concatenate([clump, far], axis=1))
    n = pos.shape[1]
    m = np.full(n, 1.0 / n)
    solver.set_state(m, pos, np.zeros_like(pos))
    a = solver.near_accel(0.05, 1e-3)
    ao = orc.near_accel(m, pos, 0.05, 1e-3)
    assert np.max(np.linalg.norm(a - ao, axis=0)) < 1e-12 * np.max(np.linalg.norm(ao, axis=0))
    assert np.all(a[:, -3:] == 0.0)


def test_near_accel_far_flung(solver, orc):
    """Positions far beyond the cell table share its edge columns: still exact,
    including a close pair sitting there and particles on the far side."""
    rng = np.random.default_rng(9)
    clump = rng.normal(0.0, 0.2, (3, 400))
    far = np.array([[3000.0, 3000.0 + 0.01, -2500.0, 0.0, 1e5],
                    [5.0, 5.0, 2500.0, -4000.0, 1e5],
                    [1.0, 1.0 + 0.02, 0.0, 7000.0, -1e5]])
    pos = np.ascontiguousarray(np.concatenate([clump, far], axis=1))
    n = pos.shape[1]
    m = np.full(n, 1.0 / n)
    solver.set_state(m, pos, np.zeros_like(pos))
    a = solver.near_accel(0.5, 1e-3)
    ao = orc.near_accel(m, pos, 0.5, 1e-3)
    assert np.max(np.linalg.norm(a - ao, axis=0)) < 1e-12 * np.max(np.linalg.norm(ao, axis=0))
    assert np.linalg.norm(ao[:, 400]) > 0 and np.all(a[:, 402:] == 0.0)


@pytest.mark.parametrize("shift", [0.0, 0.01, 0.5, 6.0])
def test_near_accel_kept_grid_after_motion(solver, orc, shift):
    """The grid of one build searched after every particle moved by up to
    shift * rc (0: no motion; 0.5: a wider column and slab reach; 6: wider
    than the scan bound at this N, so the whole array is scanned): the same
    pairs as a new grid, to FP64 rounding."""
    n, rc, eps = 3000, 0.1, 0.01
    m, pos, vel, _ = nbody_amd.init_system("plummer", n, 4)
    solver.set_state(m, pos, vel)
    solver.near_accel(rc, eps)   # builds the grid at pos
    rng = np.random.default_rng(17)
    u = rng.normal(size=(3, n))
    u *= rng.uniform(0.0, 1.0, n) ** (1.0 / 3.0) / np.linalg.norm(u, axis=0)
    moved = np.ascontiguousarray(pos + shift * rc * u)
    solver.set_state(m, moved, vel)
    a = solver.near_accel_kept(rc, eps)
    ao = orc.near_accel(m, moved, rc, eps)
    scale = np.max(np.linalg.norm(ao, axis=0))
    assert np.max(np.linalg.norm(a - ao, axis=0)) < 1e-12 * scale
    fresh = solver.near_accel(rc, eps)
    assert np.max(np.linalg.norm(a - fresh, axis=0)) < 1e-13 * scale
    if shift == 0.0:   # nothing moved: the kept grid's sums are the new grid's
        assert np.array_equal(a, fresh)


def test_near_accel_kept_mixed_after_motion(solver, orc):
    n, rc, eps = 4000, 0.1, 0.005
    m, pos, vel, _ = nbody_amd.init_system("plummer", n, 8)
    solver.set_state(m, pos, vel)
    solver.near_accel(rc, eps, form="mixed")
    moved = np.ascontiguousarray(pos + 0.05 * vel)   # |v| < 1.5: up to 0.75 rc
    solver.set_state(m, moved, vel)
    a = solver.near_accel_kept(rc, eps, form="mixed")
    ao = orc.near_accel(m, moved, rc, eps)
    scale = np.max(np.linalg.norm(ao, axis=0))
    assert np.max(np.linalg.norm(a - ao, axis=0)) < 1e-5 * scale


def test_near_accel_kept_rebuilds_for_another_radius_or_count(solver):
    """No grid of the same N and rc: a new one, as nb_near_accel builds."""
    m, pos, vel, _ = nbody_amd.init_system("plummer", 2000, 6)
    solver.set_state(m, pos, vel)
    solver.near_accel(0.2, 0.01)
    a = solver.near_accel_kept(0.1, 0.01)
    assert np.array_equal(a, solver.near_accel(0.1, 0.01))
    m, pos, vel, _ = nbody_amd.init_system("plummer", 2100, 6)
    solver.set_state(m, pos, vel)
    a = solver.near_accel_kept(0.1, 0.01)
    assert np.array_equal(a, solver.near_accel(0.1, 0.01))


def test_near_accel_rejects_bad_radius(solver):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 64, 1)
    solver.set_state(m, pos, vel)
    with pytest.raises(nbody_amd.NBodyError):
        solver.near_accel(0.0, 0.01)
    with pytest.raises(nbody_amd.NBodyError):
        solver.near_accel(float("inf"), 0.01)


def test_learned_split_matches_restatement(solver, orc):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 400, 2)
    steps, dt, eps, rc = 30, 1.0 / 256, 0.05, 0.3
    solver.learned_configure(period=4, threshold=1e-3, exact_form="fp64")
    solver.learned_set_mode(1)
    solver.learned_set_flag_ratio(2e-3)
    solver.learned_set_near_radius(rc)
    try:
        solver.set_state(m, pos, vel)
        solver.step(steps, dt, 2, "learned", eps)
        p, v = solver.get_state()
        st = solver.learned_stats()
    finally:
        solver.learned_set_near_radius(0.0)
    pr, vr, sr = restated_learned(orc, m, pos, vel, steps, dt, eps, 4, 1e-3, 2e-3, rc=rc)
    assert st["exact_evals"] == sr["exact"] and st["predicted_evals"] == sr["predicted"]
    assert st["subset_rows"] == sr["subset"]
    assert sr["predicted"] > 0
    assert np.max(np.abs(p - pr)) < 1e-11
    assert np.max(np.abs(v - vr)) < 1e-9 * np.max(np.abs(vr))


@pytest.mark.parametrize("n", [1, 2, 3, 63, 65, 130])
def test_near_accel_tiny_systems(solver, orc, n):
    """Systems smaller than a chunk, one particle, two coincident ones."""
    rng = np.random.default_rng(n)
    pos = np.ascontiguousarray(rng.uniform(-0.2, 0.2, (3, n)))
    if n >= 2:
        pos[:, 1] = pos[:, 0]   # a coincident pair contributes nothing
    m = rng.uniform(0.5, 1.5, n) / n
    solver.set_state(m, pos, np.zeros_like(pos))
    a = solver.near_accel(0.3, 0.01)
    ao = orc.near_accel(m, pos, 0.3, 0.01)
    scale = max(np.max(np.linalg.norm(ao, axis=0)), 1e-300)
    assert np.max(np.linalg.norm(a - ao, axis=0)) <= 1e-12 * scale
    if n == 1:
        assert np.all(a == 0.0)


@pytest.mark.parametrize("n", [3, 17, 100])
def test_learned_tiny_systems_match_restatement(solver, orc, n):
    from test_gpu_learned import restated_learned
    m, pos, vel, _ = nbody_amd.init_system("plummer", n, 21)
    solver.learned_configure(period=4, threshold=1e-1, exact_form="fp64")
    solver.learned_set_mode(1)
    solver.learned_set_flag_ratio(0.5)
    solver.learned_set_near_radius(0.3)
    try:
        solver.set_state(m, pos, vel)
        solver.step(20, 1.0 / 256, 2, "learned", 0.05)
        p, _ = solver.get_state()
        st = solver.learned_stats()
    finally:
        solver.learned_set_near_radius(0.0)
    pr, _, sr = restated_learned(orc, m, pos, vel, 20, 1.0 / 256, 0.05, 4, 1e-1, 0.5, rc=0.3)
    assert st["exact_evals"] == sr["exact"] and st["predicted_evals"] == sr["predicted"]
    assert np.max(np.abs(p - pr)) < 1e-11


def test_learned_split_bitwise_repeatable(solver):
    """The near field runs on a second stream beside the all-pairs kernel and
    the extrapolation: two runs from the same state agree bit for bit (N =
    20000 with the symmetric mixed kernel, rc = 0.1, both stream orders
    exercised by exact and extrapolated substeps)."""
    m, pos, vel, _ = nbody_amd.init_system("plummer", 20000, 6)
    runs = []
    try:
        for _ in range(2):
            solver.learned_configure(period=4, threshold=3e-2, exact_form="mixed")   # counters start at 0
            solver.learned_set_mode(1)
            solver.learned_set_near_radius(0.1)
            solver.set_state(m, pos, vel)
            solver.step(24, 1.0 / 256, 2, "learned", 0.01)
            runs.append((*solver.get_state(), solver.learned_stats()))
    finally:
        solver.learned_set_near_radius(0.0)
    assert runs[0][2]["predicted_evals"] > 0
    assert runs[0][2] == runs[1][2]
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
