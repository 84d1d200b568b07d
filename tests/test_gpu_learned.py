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


def test_mlp_forward_single_path(solver, orc):
    """The hand-built single-path network of the oracle's closed-form test
    through the MFMA forward: same layout, biases and SiLU (bf16 operands)."""
    from test_oracle_learned import single_path_expect, single_path_params
    P = single_path_params(orc)
    solver.mlp_set_params(P.astype(np.float32))
    x = np.random.default_rng(9).uniform(-2, 2, (40, 15)).astype(np.float32)
    y = solver.mlp_forward(x)
    np.testing.assert_allclose(y, single_path_expect(x.astype(np.float64)), rtol=0, atol=2e-2)
    yo = orc.mlp_forward(P, x.astype(np.float64), bf16=True)
    assert np.max(np.abs(y - yo)) < 1e-5


def test_adam_steps_vs_oracle(solver, orc):
    """Three GPU Adam steps with different gradients against the oracle's
    Adam (FP32 arithmetic on the device, FP64 in the oracle)."""
    p0 = random_params(8)
    solver.mlp_set_params(p0)
    rng = np.random.default_rng(10)
    grads = [rng.standard_normal(p0.shape[0]).astype(np.float32) * s for s in (1.0, 0.3, 2.0)]
    grads[1][:100] = 0.0   # zero gradient after a non-zero one: the moments alone move the weights
    p, m, v = p0.astype(np.float64), np.zeros(p0.shape[0]), np.zeros(p0.shape[0])
    for t, g in enumerate(grads, start=1):
        solver.mlp_adam_step(g, 1e-3)
        orc.adam(p, g.astype(np.float64), m, v, 1e-3, t)
    pg = solver.mlp_get_params().astype(np.float64)
    step = np.abs(p - p0)
    assert np.median(step) > 1e-4       # the parameters moved
    assert np.max(np.abs(pg - p)) < 1e-6 * np.max(np.abs(p0)) + 1e-4 * np.max(step)
    # the moments restart with new parameters
    solver.mlp_set_params(p0)
    solver.mlp_adam_step(grads[0], 1e-3)
    q, mq, vq = p0.astype(np.float64), np.zeros(p0.shape[0]), np.zeros(p0.shape[0])
    orc.adam(q, grads[0].astype(np.float64), mq, vq, 1e-3, 1)
    assert np.max(np.abs(solver.mlp_get_params() - q)) < 1e-6


# --- learned stepping ----------------------------------------------------------

def restated_learned(orc, m, pos, vel, steps, dt, eps, period, threshold, rho, params=None, tau=1.0, rc=0.0,
                     kappa=1e-3, rough_share=0.25):
    """numpy restatement of the learned form's schedule (order-2 DKD): the
    history holds the last four exact samples with their substeps, an
    extrapolated substep takes the cubic through them (Lagrange weights from
    the oracle), plus the network scaled by the nodal polynomial when the
    samples are spaced one period apart (its outputs are components in each
    particle's frame from the oracle) and the network passed the last
    check (checks test the cubic alone and, where it applies, the network).
    With rc > 0 the history holds
    a - a_near, the near field is summed exactly on every substep and added to
    the extrapolated remainder.  Particles rough, or held after missing the
    last check by more than kappa a_rms, are summed exactly.  A substep due
    for extrapolation right after one that extrapolated with more than
    rough_share N such rows runs the direct sum instead, and so do the
    substeps of the next 8 periods (without checks), twice as many for each
    further time in a row, up to 64 periods."""
    pos, vel = pos.copy(), vel.copy()
    n = m.shape[0]
    hist, htime = [], []
    since, fallback, sub, net_ok, hold, prev, streak = 0, True, 0, False, 0, -1, 0
    held = np.zeros(n, dtype=bool)
    st = {"exact": 0, "predicted": 0, "subset": 0, "rough": 0}

    def predict(allow_net):
        t = np.array(htime[::-1], dtype=float)       # newest first
        w = orc.lagrange4(t, float(sub))
        a0, a1, a2, a3 = (hist[-1].T, hist[-2].T, hist[-3].T, hist[-4].T)
        x, s, _, fr = orc.features(a0, a1, a2, a3, vel.T)
        cub = w[0] * hist[-1] + w[1] * hist[-2] + w[2] * hist[-3] + w[3] * hist[-4]
        even = all(t[k] - t[k + 1] == period for k in range(3))
        used = params is not None and even and allow_net
        pred = cub
        if used:
            wnet = np.prod(sub - t) / np.prod(t[0] + period - t)
            y = orc.mlp_forward(params, x, bf16=True)
            corr = np.einsum("ik,ikj->ij", y, fr)          # components in each particle's frame
            corr = (s * float(np.float32(tau * wnet)))[:, None] * corr
            vt = vel.T
            v2 = np.sum(vt ** 2, axis=1)
            pr = np.where(v2 > 0, np.sum(corr * vt, axis=1) / np.where(v2 > 0, v2, 1.0), 0.0)
            corr = corr - pr[:, None] * vt                    # no work: the part along v is removed
            pred = cub + corr.T
        rough = ~(s <= rho * np.linalg.norm(a0, axis=1)) | held
        return cub, pred, used, rough

    pos += 0.5 * dt * vel
    for t in range(steps):
        full = len(hist) == 4
        held_exact = hold > 0
        hold = max(hold - 1, 0)
        exact = (not full) or held_exact or fallback or since + 1 >= period
        near = orc.near_accel(m, pos, rc, eps) if rc > 0 else np.zeros_like(pos)
        prev_rough, prev = prev, -1
        rough_prev = prev_rough >= 0 and float(prev_rough) > rough_share * float(n)
        if prev_rough >= 0 and not rough_prev:
            streak = 0
        if not exact and rough_prev:
            exact, held_exact, hold = True, True, (8 * period) << min(streak, 3)
            streak += 1
            st["rough"] += 1
        if exact:
            acc = orc.accel(m, pos, eps)
            st["exact"] += 1
            if full and not held_exact:
                cub, pred, used, rough = predict(True)
                smooth = ~rough
                a2 = np.sum(acc[:, smooth] ** 2)
                rel = lambda d: np.sqrt(np.sum(d[:, smooth] ** 2) / a2) if a2 > 0 else np.inf
                err_c = rel((acc - near) - cub)
                err_n = rel((acc - near) - pred) if used else np.inf
                net_ok = used and err_n <= threshold
                err = err_n if net_ok else err_c
                fallback = not (err <= threshold)
                D = (acc - near) - (pred if net_ok else cub)
                held = ~(np.sum(D ** 2, axis=0) <= kappa ** 2 * a2 / n)
            hist.append(acc - near)
            htime.append(sub)
            hist, htime = hist[-4:], htime[-4:]
            since = 0
        else:
            _, pred, _, rough = predict(net_ok)
            acc = pred + near
            if rough.any():
                ex = orc.accel(m, pos, eps)
                acc[:, rough] = ex[:, rough]
            st["predicted"] += 1
            st["subset"] += int(rough.sum())
            prev = int(rough.sum())
            since += 1
        sub += 1
        vel += dt * acc
        pos += 0.5 * dt * vel
        if t + 1 < steps:
            pos += 0.5 * dt * vel
    return pos, vel, st


@pytest.mark.parametrize("rc", [0.0, 0.3])
def test_rough_guard_runs_the_direct_sum(solver, rc):
    """More than N/4 exact rows on an extrapolated substep (here nearly every
    particle is held after the first check: kappa = 1e-12): the next substep
    runs the direct sum instead, and so do the next 8 periods' substeps, then
    16 periods' the second time in a row.  Substep 4 checks; 5 and 39
    extrapolate (nearly every row summed exactly); 6 and 40 are the guard's;
    the rest are exact.  With a near-field split (rc = 0.3) the held
    substeps that cannot reach the history skip the near field."""
    m, pos, vel, _ = nbody_amd.init_system("plummer", 2000, 9)
    steps, dt, eps = 80, 1.0 / 256, 0.01
    solver.set_state(m, pos, vel)
    solver.step(steps, dt, 2, "fp64", eps)
    p_exact, v_exact = solver.get_state()
    solver.learned_configure(period=4, threshold=1e-2, exact_form="fp64")
    solver.learned_set_mode(1)
    solver.learned_set_near_radius(rc)
    solver.learned_set_flag_ratio(0.1)
    solver.learned_set_hold_ratio(1e-12)
    try:
        solver.set_state(m, pos, vel)
        solver.step(steps, dt, 2, "learned", eps)
        p, v = solver.get_state()
        st = solver.learned_stats()
    finally:
        solver.learned_set_hold_ratio(1e-3)
        solver.learned_set_near_radius(0.0)
    assert st["checks"] == 1 and st["fallbacks"] == 0
    assert st["rough_exact"] == 2 and st["predicted_evals"] == 2 and st["exact_evals"] == steps - 2
    # (all but the far, smoothest particles miss by more than 1e-12 a_rms)
    assert 2 * m.shape[0] // 4 < st["subset_rows"] <= 2 * m.shape[0]
    # every force was summed exactly (three substeps row by row): the exact run to rounding
    assert np.max(np.abs(p - p_exact)) < 1e-12 * np.max(np.abs(p_exact))
    assert np.max(np.abs(v - v_exact)) < 1e-10 * np.max(np.abs(v_exact))


@pytest.mark.parametrize("n", [1, 2, 3, 65])
@pytest.mark.parametrize("rc", [0.0, 0.3])
def test_learned_tiny_systems(solver, orc, n, rc):
    """Systems smaller than a wave: the schedule, the device-sized exact rows
    and the near field still follow the restatement (one particle: nothing to
    check, every substep exact)."""
    m, pos, vel, _ = nbody_amd.init_system("plummer", max(n, 2), 5)
    m, pos, vel = m[:n] * 0 + 1.0 / n, np.ascontiguousarray(pos[:, :n]), np.ascontiguousarray(vel[:, :n])
    steps, dt, eps = 24, 1.0 / 256, 0.05
    solver.learned_configure(period=4, threshold=1e-2, exact_form="fp64")
    solver.learned_set_mode(1)
    solver.learned_set_flag_ratio(0.1)
    solver.learned_set_near_radius(rc)
    try:
        solver.set_state(m, pos, vel)
        solver.step(steps, dt, 2, "learned", eps)
        p, v = solver.get_state()
        st = solver.learned_stats()
    finally:
        solver.learned_set_near_radius(0.0)
    pr, vr, sr = restated_learned(orc, m, pos, vel, steps, dt, eps, 4, 1e-2, 0.1, rc=rc)
    assert st["exact_evals"] == sr["exact"] and st["predicted_evals"] == sr["predicted"]
    assert st["subset_rows"] == sr["subset"] and st["rough_exact"] == sr["rough"]
    assert np.all(np.isfinite(p)) and np.all(np.isfinite(v))
    assert np.max(np.abs(p - pr)) < 1e-11 * max(1.0, np.max(np.abs(pr)))
    if n == 1:
        assert st["predicted_evals"] == 0


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


def test_corrupted_network_never_used(solver):
    """NaN weights fail every check they take part in, so the network is never
    applied: the run equals the cubic-only run bit for bit."""
    m, pos, vel, _ = nbody_amd.init_system("plummer", 1000, 4)
    runs = {}
    for mode in (1, 0):
        solver.learned_configure(period=4, threshold=1e-2, exact_form="mixed")
        solver.learned_set_mode(mode)
        if mode == 0:
            solver.mlp_set_params(np.full(nbody_amd.mlp_param_count(), np.nan, dtype=np.float32))
        solver.set_state(m, pos, vel)
        solver.step(40, 1.0 / 256, 2, "learned", 0.01)
        runs[mode] = (solver.get_state()[0], solver.learned_stats())
    assert runs[0][1]["predicted_evals"] > 0
    assert runs[0][1]["predicted_evals"] == runs[1][1]["predicted_evals"]
    assert np.all(np.isfinite(runs[0][0]))
    assert np.array_equal(runs[0][0], runs[1][0])


def test_training_and_gain(solver, capsys):
    """Collect exact trajectories with the near-field split, train, and
    measure what the network gains: every check evaluates the cubic alone and
    the cubic plus the network on the same history, so the RMS errors compare
    pair by pair.  On a fresh system the network is validated, used, and
    beats the cubic alone by a margin a zero-output network (gain exactly 1)
    cannot reach."""
    eps, dt, P, rc = 0.005, 1.0 / 256, 4, 0.3
    m, pos, vel, _ = nbody_amd.init_system("plummer", 32768, 11)
    solver.learned_configure(period=P, threshold=1e-2, exact_form="mixed")
    solver.learned_set_near_radius(rc)
    solver.learned_set_flag_ratio(0.03)
    try:
        solver.learned_reset(7)
        solver.set_state(m, pos, vel)
        solver.learned_collect(100, dt, 2, "mixed", eps)
        loss0 = solver.learned_train(1, 4096, 1e-3, 7)
        loss1 = solver.learned_train(3000, 4096, 1e-3, 7)
        assert np.isfinite(loss1) and loss1 < loss0
        m2, pos2, vel2, _ = nbody_amd.init_system("plummer", 32768, 12)
        solver.learned_configure(period=P, threshold=1e-2, exact_form="mixed")
        solver.learned_set_mode(0)
        solver.set_state(m2, pos2, vel2)
        gains = []
        for _ in range(16):   # 16 cycles; one check each once the history is full
            solver.step(P, dt, 2, "learned", eps)
            st = solver.learned_stats()
            if st["last_check_err"] > 0:
                gains.append(st["last_cubic_err"] / st["last_check_err"])
        st = solver.learned_stats()
    finally:
        solver.learned_set_near_radius(0.0)
    with capsys.disabled():
        print(f"\n[learned] loss {loss0:.3e} -> {loss1:.3e}; check RMS error cubic / network over the last 8 "
              f"checks: {np.round(gains[-8:], 3).tolist()}; network used on {st['net_evals']} of "
              f"{st['predicted_evals']} extrapolated substeps")
    assert st["net_evals"] > 0
    # a network that outputs zeros has a gain of exactly 1 at every check
    assert st["net_checks"] >= 8
    assert st["net_cubic_err_sum"] / st["net_err_sum"] >= 1.03
    assert np.median(gains[-8:]) >= 1.03


def test_network_mode_matches_restatement(solver, orc):
    """Network on: its output enters scaled by |D3| tau and the nodal-polynomial
    weight, and only on histories spaced one period apart."""
    m, pos, vel, _ = nbody_amd.init_system("plummer", 300, 6)
    steps, dt, eps = 40, 1.0 / 256, 0.05
    p = random_params(21, scale=0.3)
    solver.learned_reset(3)
    solver.mlp_set_params(p)
    solver.learned_configure(period=4, threshold=1e-2, exact_form="fp64")
    solver.learned_set_mode(0)
    solver.learned_set_flag_ratio(0.5)
    solver.set_state(m, pos, vel)
    solver.step(steps, dt, 2, "learned", eps)
    pg, vg = solver.get_state()
    st = solver.learned_stats()
    pr, vr, sr = restated_learned(orc, m, pos, vel, steps, dt, eps, 4, 1e-2, 0.5, params=p.astype(np.float64))
    assert st["exact_evals"] == sr["exact"] and st["predicted_evals"] == sr["predicted"]
    assert sr["predicted"] >= 20
    assert np.max(np.abs(pg - pr)) < 1e-10
    assert np.max(np.abs(vg - vr)) < 1e-8 * np.max(np.abs(vr))
    # and the network changed the trajectory (against the cubic alone)
    pc, _, _ = restated_learned(orc, m, pos, vel, steps, dt, eps, 4, 1e-2, 0.5)
    assert np.max(np.abs(pc - pr)) > 1e-12


def test_learned_fourth_order_uses_real_substep_times(solver):
    """Order-4 Yoshida substeps are unevenly spaced in time; the learned form
    extrapolates in the positions' time and stays close to the exact run."""
    m, pos, vel, _ = nbody_amd.init_system("plummer", 4096, 13)
    eps, dt, steps = 0.01, 1.0 / 128, 40
    solver.set_state(m, pos, vel)
    solver.step(steps, dt, 4, "fp64", eps)
    p_exact, _ = solver.get_state()
    solver.learned_configure(period=4, threshold=1e-2, exact_form="fp64")
    solver.learned_set_mode(1)
    solver.learned_set_near_radius(0.3)
    solver.learned_set_rough_share(1.0)   # extrapolate however many rows are exact (dt = 1/128 flags many)
    try:
        solver.set_state(m, pos, vel)
        solver.step(steps, dt, 4, "learned", eps)
        p, _ = solver.get_state()
        st = solver.learned_stats()
    finally:
        solver.learned_set_near_radius(0.0)
        solver.learned_set_rough_share(0.25)
    assert st["predicted_evals"] > 40 and st["fallbacks"] <= 2, st
    assert st["max_check_err"] < 1e-2
    assert np.max(np.abs(p - p_exact)) < 1e-6 * np.max(np.abs(p_exact))


def test_learned_new_dt_restarts_history(solver):
    m, pos, vel, _ = nbody_amd.init_system("plummer", 2048, 14)
    solver.learned_configure(period=4, threshold=1e-2, exact_form="fp64")
    solver.learned_set_mode(1)
    solver.set_state(m, pos, vel)
    solver.step(12, 1.0 / 256, 2, "learned", 0.01)
    st1 = solver.learned_stats()
    solver.step(6, 1.0 / 512, 2, "learned", 0.01)   # the history restarts: 4 exact samples first
    st2 = solver.learned_stats()
    assert st2["exact_evals"] - st1["exact_evals"] >= 5
