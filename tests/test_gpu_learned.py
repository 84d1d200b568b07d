"""The learned extrapolator on the GPU: the MFMA network kernels against the
CPU oracle (forward and gradient, bf16 operands emulated there), the learned
stepping against a numpy restatement of its schedule, the fallback rule, and
what the trained network gains over the cubic extrapolation alone."""
import numpy as np
import pytest

import nbody_amd

pytestmark = pytest.mark.gpu


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    sizes = [(128, 15), (128,), (128, 128), (128,), (128, 128), (128,), (3, 128), (3,)]
    parts = []
    for s in sizes:
        fan_in = s[1] if len(s) == 2 else 8
        parts.append(rng.uniform(-1, 1, int(np.prod(s))) * scale / np.sqrt(fan_in))
    return np.concatenate(parts).astype(np.float32)


SLICES = {"W1": (0, 1920), "b1": (1920, 2048), "W2": (2048, 18432), "b2": (18432, 18560), "W3": (18560, 34944),
          "b3": (34944, 35072), "W4": (35072, 35456), "b4": (35456, 35459)}


@pytest.mark.parametrize("rows", [16, 37, 1000])
def test_mlp_forward_vs_oracle(solver, orc, rows):
    p = random_params(1)
    solver.mlp_set_params(p)
    x = np.random.default_rng(rows).uniform(-1, 1, (rows, 15)).astype(np.float32)
    y = solver.mlp_forward(x)
    yo = orc.mlp_forward(p.astype(np.float64), x.astype(np.float64), bf16=True)
    assert np.max(np.abs(y - yo)) < 2e-3 * np.max(np.abs(yo))
    # and close to the unrounded FP64 network (bf16 operands cost ~1e-2)
    y64 = orc.mlp_forward(p.astype(np.float64), x.astype(np.float64), bf16=False)
    assert np.max(np.abs(y - y64)) < 5e-2 * np.max(np.abs(y64))


def test_mlp_gradient_vs_oracle(solver, orc):
    p = random_params(2)
    solver.mlp_set_params(p)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, (200, 15)).astype(np.float32)
    t = rng.uniform(-1, 1, (200, 3)).astype(np.float32)
    g, loss = solver.mlp_grad(x, t)
    go, lo = orc.mlp_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), bf16=True)
    assert abs(loss - lo) < 1e-3 * abs(lo)
    for name, (a, b) in SLICES.items():
        err = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        assert err < 2e-2, (name, err)


def test_mlp_gradient_descends(solver):
    # one plain gradient step lowers the loss (sign and scale of the GPU gradient)
    p = random_params(4)
    solver.mlp_set_params(p)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (256, 15)).astype(np.float32)
    t = rng.uniform(-0.5, 0.5, (256, 3)).astype(np.float32)
    g, l0 = solver.mlp_grad(x, t)
    solver.mlp_set_params((p - 0.05 * g / np.linalg.norm(g)).astype(np.float32))
    _, l1 = solver.mlp_grad(x, t)
    assert l1 < l0


# --- learned stepping ----------------------------------------------------------

def restated_learned(orc, m, pos, vel, steps, dt, eps, period, threshold, rho, params=None, tau=1.0):
    """numpy restatement of the learned form's schedule (order-2 DKD)."""
    pos, vel = pos.copy(), vel.copy()
    hist = []
    since, fallback = 0, True
    st = {"exact": 0, "predicted": 0, "subset": 0}

    def predict():
        a0, a1, a2, a3 = (hist[-1].T, hist[-2].T, hist[-3].T, hist[-4].T)
        x, s, cub = orc.features(a0, a1, a2, a3, vel.T)
        if params is not None:
            y = orc.mlp_forward(params, x, bf16=True)
            cub = cub + (s * tau)[:, None] * y
        rough = ~(s <= rho * np.linalg.norm(a0, axis=1))
        return cub.T, rough

    pos += 0.5 * dt * vel
    for t in range(steps):
        full = len(hist) == 4
        exact = (not full) or fallback or since + 1 >= period
        if exact:
            acc = orc.accel(m, pos, eps)
            st["exact"] += 1
            if full:
                pred, rough = predict()
                smooth = ~rough
                e2 = np.sum((pred - acc)[:, smooth] ** 2)
                a2 = np.sum(acc[:, smooth] ** 2)
                fallback = not (np.sqrt(e2 / a2) <= threshold)
            since = 0
        else:
            acc, rough = predict()
            if rough.any():
                ex = orc.accel(m, pos, eps)
                acc[:, rough] = ex[:, rough]
            st["predicted"] += 1
            st["subset"] += int(rough.sum())
            since += 1
        hist.append(acc.copy())
        hist = hist[-4:]
        vel += dt * acc
        pos += 0.5 * dt * vel
        if t + 1 < steps:
            pos += 0.5 * dt * vel
    return pos, vel, st


def test_threshold_zero_is_all_exact_bitwise(solver):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 2000, 9)
    solver.set_state(m, pos, vel)
    solver.step(24, 1.0 / 256, 2, "fp64", 0.01)
    p_exact, v_exact = solver.get_state()
    solver.learned_configure(period=4, threshold=0.0, exact_form="fp64")
    solver.set_state(m, pos, vel)
    solver.step(24, 1.0 / 256, 2, "learned", 0.01)
    p, v = solver.get_state()
    st = solver.learned_stats()
    assert st["predicted_evals"] == 0 and st["exact_evals"] == 24
    assert np.array_equal(p, p_exact) and np.array_equal(v, v_exact)


def test_cubic_mode_matches_restatement(solver, orc):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 400, 2)
    steps, dt, eps = 30, 1.0 / 256, 0.05
    solver.learned_configure(period=4, threshold=1e-3, exact_form="fp64")
    solver.learned_set_mode(1)
    solver.learned_set_flag_ratio(2e-3)
    solver.set_state(m, pos, vel)
    solver.step(steps, dt, 2, "learned", eps)
    p, v = solver.get_state()
    st = solver.learned_stats()
    pr, vr, sr = restated_learned(orc, m, pos, vel, steps, dt, eps, 4, 1e-3, 2e-3)
    assert st["exact_evals"] == sr["exact"] and st["predicted_evals"] == sr["predicted"]
    assert st["subset_rows"] == sr["subset"]
    assert sr["predicted"] > 0 and sr["subset"] > 0
    assert np.max(np.abs(p - pr)) < 1e-11
    assert np.max(np.abs(v - vr)) < 1e-9 * np.max(np.abs(vr))


def test_corrupted_network_falls_back(solver):
    # NaN weights make every check fail: the run equals the all-exact run
    m, pos, vel, _ = nbody_amd.init_system("plummer", 1000, 4)
    solver.set_state(m, pos, vel)
    solver.step(20, 1.0 / 256, 2, "mixed", 0.01)
    p_exact, _ = solver.get_state()
    solver.learned_configure(period=4, threshold=1e-2, exact_form="mixed")
    solver.learned_set_mode(0)
    solver.mlp_set_params(np.full(nbody_amd.mlp_param_count(), np.nan, dtype=np.float32))
    solver.set_state(m, pos, vel)
    solver.step(20, 1.0 / 256, 2, "learned", 0.01)
    p, _ = solver.get_state()
    st = solver.learned_stats()
    assert st["predicted_evals"] == 0 and st["fallbacks"] == 20 - 4
    assert np.array_equal(p, p_exact)


def test_training_and_gain(solver, capsys):
    """Collect exact trajectories, train, and compare the network-corrected
    extrapolation with the cubic alone on a fresh system."""
    eps, dt = 0.005, 1.0 / 256
    m, pos, vel, _ = nbody_amd.init_system("plummer", 16384, 11)
    solver.learned_configure(period=4, threshold=1e-2, exact_form="mixed")
    solver.learned_reset(7)
    solver.set_state(m, pos, vel)
    solver.learned_collect(64, dt, 2, "mixed", eps)
    loss0 = solver.learned_train(1, 4096, 1e-3, 7)
    loss1 = solver.learned_train(600, 4096, 1e-3, 7)
    assert np.isfinite(loss1) and loss1 < loss0
    # fresh system, same statistics: cubic alone vs cubic + network
    m2, pos2, vel2, _ = nbody_amd.init_system("plummer", 16384, 12)
    res = {}
    for mode in (1, 0):
        solver.learned_configure(period=4, threshold=1e-2, exact_form="mixed")
        solver.learned_set_mode(mode)
        solver.set_state(m2, pos2, vel2)
        solver.step(64, dt, 2, "learned", eps)
        res[mode] = solver.learned_stats()
    with capsys.disabled():
        print(f"\n[learned] loss {loss0:.3e} -> {loss1:.3e}; cubic-only check err max {res[1]['max_check_err']:.3e}, "
              f"with network {res[0]['max_check_err']:.3e}; fallbacks {res[1]['fallbacks']} vs {res[0]['fallbacks']}")
    assert res[0]["predicted_evals"] > 0
    assert np.isfinite(res[0]["max_check_err"])
