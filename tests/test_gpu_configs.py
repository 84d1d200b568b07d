"""The workload configurations at their stated sizes on the GPU.

C2  Plummer N = 32768 (a = 1, r < 20), eps = 0.01, dt = 1/128, 1000 steps:
    FP64 and mixed forces against the oracle on sampled rows, two leapfrog
    steps against the oracle's, and the 1000-step runs' energy drift and the
    gap between the FP64 and mixed trajectories.  The reference's own
    measurement at this N is its single-node run
    (final_results/single_node_combined.csv:18).
C4  the network trained on 200 exact steps of a Plummer N = 65536 with the
    learned form's production settings: the loss falls and the network passes
    its checks on a system it was not trained on.
C5  the learned form at N = 262144 (eps = 0.005, dt = 1/256) with that
    network: no fallback, checks within the threshold, the network in use,
    the energy bound, a trajectory close to the mixed form's; and with
    threshold 0 (every substep exact) the mixed form's run bit for bit.
"""
import numpy as np
import pytest

import nbody_amd
from conftest import rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 2e-5
# the learned form's production settings (bench.py LEARNED)
LEARNED = {"period": 4, "threshold": 3e-3, "rc": 0.25, "rho": 0.1, "kappa": 3e-3}


def dE(solver, eps, E0):
    T, U = solver.energy(eps)
    return abs((T + U - E0) / E0)


@pytest.fixture(scope="module")
def s():
    sv = nbody_amd.Solver(0)
    yield sv
    sv.close()


@pytest.fixture(scope="module")
def c2():
    m, pos, vel, _ = nbody_amd.init_system("plummer", 32768, 2, [1.0, 20.0])
    return m, pos, vel


def test_c2_forces_vs_oracle(s, orc, c2):
    m, pos, vel = c2
    n = m.shape[0]
    s.set_state(m, pos, vel)
    rows = np.random.default_rng(2).choice(n, 1024, replace=False)
    pj = np.stack([pos[0], pos[1], pos[2], m], axis=1)
    ref = orc.accel_block(pj[rows], pj, 0.01).T
    assert rel_err(s.accel("fp64", 0.01)[:, rows], ref) < FP64_TOL
    assert rel_err(s.accel("mixed", 0.01)[:, rows], ref) < MIXED_TOL
    assert s.kernel_info()["kernel"] == "symmetric"


def test_c2_leapfrog_vs_oracle(s, orc, c2):
    m, pos, vel = c2
    s.set_state(m, pos, vel)
    s.step(2, 1.0 / 128, 2, "fp64", 0.01)
    p, v = s.get_state()
    po, vo = orc.leapfrog(m, pos, vel, 2, 1.0 / 128, 2, 0.01)
    assert np.max(np.abs(p - po)) < 1e-12 * np.max(np.abs(po))
    assert np.max(np.abs(v - vo)) < 1e-11 * np.max(np.abs(vo))


def test_c2_thousand_steps_fp64_and_mixed(s, c2, capsys):
    m, pos, vel = c2
    eps, dt = 0.01, 1.0 / 128
    out = {}
    for form in ("fp64", "mixed"):
        s.set_state(m, pos, vel)
        T0, U0 = s.energy(eps)
        s.step(1000, dt, 2, form, eps)
        out[form] = (dE(s, eps, T0 + U0),) + s.get_state()
    gap = np.linalg.norm(out["mixed"][1] - out["fp64"][1], axis=0)
    with capsys.disabled():
        print(f"\n[C2] |dE/E| over 1000 steps: fp64 {out['fp64'][0]:.2e}, mixed {out['mixed'][0]:.2e}; "
              f"position gap mixed - fp64: median {np.median(gap):.2e}, 99th percentile "
              f"{np.quantile(gap, 0.99):.2e}, max {gap.max():.2e}")
    assert out["fp64"][0] <= 1e-6 and out["mixed"][0] <= 1e-6
    # the FP32 pair arithmetic perturbs every force by ~1e-7 relative; close
    # encounters amplify it, but the bulk of the system stays on the FP64 track
    assert np.median(gap) < 1e-6
    assert np.quantile(gap, 0.99) < 1e-5


@pytest.fixture(scope="module")
def trained(s):
    """C4: 200 exact steps of a Plummer N = 65536 with the production
    settings, then the bench's training schedule."""
    m, pos, vel, _ = nbody_amd.init_system("plummer", 65536, 2, [1.0, 20.0])
    s.learned_configure(period=LEARNED["period"], threshold=LEARNED["threshold"], exact_form="mixed")
    s.learned_set_near_radius(LEARNED["rc"])
    s.learned_set_flag_ratio(LEARNED["rho"])
    s.learned_set_hold_ratio(LEARNED["kappa"])
    s.learned_reset(1)
    s.set_state(m, pos, vel)
    s.learned_collect(200, 1.0 / 256, 2, "mixed", 0.005)
    loss0 = s.learned_train(1, 4096, 1e-3, 1)
    loss1 = s.learned_train(4000, 4096, 1e-3, 1)
    loss2 = s.learned_train(4000, 4096, 2e-4, 2)
    return loss0, loss1, loss2


def configure(s, mode, threshold=None):
    s.learned_configure(period=LEARNED["period"], threshold=LEARNED["threshold"] if threshold is None else threshold,
                        exact_form="mixed")
    s.learned_set_near_radius(LEARNED["rc"])
    s.learned_set_flag_ratio(LEARNED["rho"])
    s.learned_set_hold_ratio(LEARNED["kappa"])
    s.learned_set_mode(mode)


def test_c4_training_validates_on_a_new_system(s, trained, capsys):
    loss0, loss1, loss2 = trained
    assert np.isfinite(loss2) and loss1 < loss0 and loss2 < 0.95 * loss0
    m, pos, vel, _ = nbody_amd.init_system("plummer", 65536, 3, [1.0, 20.0])
    configure(s, 0)
    s.set_state(m, pos, vel)
    s.step(96, 1.0 / 256, 2, "learned", 0.005)
    st = s.learned_stats()
    with capsys.disabled():
        print(f"\n[C4] loss {loss0:.3f} -> {loss1:.3f} -> {loss2:.3f}; on a new N = 65536 system: "
              f"{st['net_checks']} checks with the network, mean error ratio cubic / network "
              f"{st['net_cubic_err_sum'] / max(st['net_err_sum'], 1e-300):.3f}, network on {st['net_evals']} of "
              f"{st['predicted_evals']} extrapolated substeps")
    assert st["fallbacks"] == 0
    assert st["net_checks"] >= 8 and st["net_evals"] >= 0.7 * st["predicted_evals"]
    assert st["max_check_err"] <= LEARNED["threshold"]


@pytest.fixture(scope="module")
def c5():
    m, pos, vel, _ = nbody_amd.init_system("plummer", 262144, 5, [1.0, 20.0])
    return m, pos, vel


def test_c5_learned_leg(s, trained, c5, capsys):
    m, pos, vel = c5
    eps, dt, steps = 0.005, 1.0 / 256, 96
    s.set_state(m, pos, vel)
    s.step(steps, dt, 2, "mixed", eps)
    p_mixed, _ = s.get_state()
    configure(s, 0)
    s.set_state(m, pos, vel)
    T0, U0 = s.energy(eps)
    s.step(steps, dt, 2, "learned", eps)
    drift = dE(s, eps, T0 + U0)
    p, _ = s.get_state()
    st = s.learned_stats()
    gap = np.linalg.norm(p - p_mixed, axis=0)
    with capsys.disabled():
        print(f"\n[C5] learned, {steps} steps: |dE/E| {drift:.2e}; exact {st['exact_evals']}, extrapolated "
              f"{st['predicted_evals']} (network on {st['net_evals']}), rows summed exactly {st['subset_rows']}; "
              f"check error max {st['max_check_err']:.2e}; position gap to the mixed run: median "
              f"{np.median(gap):.2e}, max {gap.max():.2e}")
    assert st["fallbacks"] == 0
    assert st["max_check_err"] <= LEARNED["threshold"]
    # the 4 samples of the cold start, the first check, then one in every period
    assert st["exact_evals"] == 4 + (steps - 4) // LEARNED["period"]
    assert st["net_evals"] >= 0.7 * st["predicted_evals"]
    assert drift <= 1e-6
    assert np.median(gap) < 1e-6


def test_c5_threshold_zero_is_the_mixed_run(s, trained, c5):
    """Every substep exact: the learned form's exact sums are the mixed
    kernel's, so the run equals the mixed form's bit for bit."""
    m, pos, vel = c5
    eps, dt, steps = 0.005, 1.0 / 256, 8
    s.set_state(m, pos, vel)
    s.step(steps, dt, 2, "mixed", eps)
    p_mixed, v_mixed = s.get_state()
    configure(s, 0, threshold=0.0)
    s.set_state(m, pos, vel)
    s.step(steps, dt, 2, "learned", eps)
    p, v = s.get_state()
    st = s.learned_stats()
    assert st["exact_evals"] == steps and st["predicted_evals"] == 0
    assert np.array_equal(p, p_mixed) and np.array_equal(v, v_mixed)
