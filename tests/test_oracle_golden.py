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


@pytest.mark.parametrize("name", ["ringround_64_p2.npz", "ringround_300_p3.npz"])
def test_oracle_ring_round_vs_reference_fixtures(orc, name):
    """The oracle's ring round against compute_forces_avx2_blocked's and
    compute_forces_avx512_rsqrt_4i_blocked's outputs for every (rank,
    recv_round): pair selection, mass indexing, signs and accumulation."""
    g = golden(name)
    P, n = g["force0"].shape[0], g["force0"].shape[3]
    blocks = [np.ascontiguousarray(g["pos"][:, r * n:(r + 1) * n]) for r in range(P)]
    for rank in range(P):
        for rr in range(P):
            f = np.ascontiguousarray(g["force0"][rank, rr].copy())
            r = np.ascontiguousarray(g["recv0"][rank, rr].copy())
            orc.ring_round(g["mass"], rank, P, rr, blocks[rank], blocks[(rank + rr) % P], f, r)
            for variant in ("avx2", "avx512"):
                for out, key in ((f, "force"), (r, "recv")):
                    ref = g[f"{key}_{variant}"][rank, rr]
                    scale = np.max(np.abs(ref - g[f"{key}0"][rank, rr]))
                    assert np.max(np.abs(out - ref)) <= 1e-12 * scale, (rank, rr, variant, key)


def ring_sum(orc, m, pos, P, n, round_fn):
    blocks = [np.ascontiguousarray(pos[:, r * n:(r + 1) * n]) for r in range(P)]
    F = np.zeros((3, P * n))
    for rank in range(P):
        f = np.zeros((3, n))
        for rr in range(P):
            partner = (rank + rr) % P
            rf = np.zeros((3, n))
            round_fn(m, rank, P, rr, blocks[rank], blocks[partner], f, rf)
            F[:, partner * n:(partner + 1) * n] += rf
        F[:, rank * n:(rank + 1) * n] += f
    return F


@pytest.mark.parametrize("P,n", [(2, 5), (3, 7), (4, 16)])
def test_oracle_ring_covers_every_pair_once(orc, P, n):
    """Over a whole ring (every rank, every round, received forces returned
    to their owners) each pair is evaluated exactly once: the sum equals the
    whole system's reduced force.  A pair rule of q > k in every round misses
    the pairs with q = k across blocks and fails."""
    rng = np.random.default_rng(P * 100 + n)
    m = rng.uniform(0.5, 1.5, P * n)
    pos = rng.standard_normal((3, P * n))
    full = orc.forces_reduced(m, pos, 0.0)
    F = ring_sum(orc, m, pos, P, n, orc.ring_round)
    assert np.max(np.abs(F - full)) <= 1e-13 * np.max(np.abs(full))

    def strict(mm, rank, PP, rr, lp, rp, f, rf):   # planted mistake: always q > k
        nn = lp.shape[1]
        partner = (rank + rr) % PP
        for k in range(nn):
            for q in range(k + 1, nn):
                d = lp[:, k] - rp[:, q]
                fk = -mm[rank * nn + k] * mm[partner * nn + q] * d / np.linalg.norm(d) ** 3
                f[:, k] += fk
                rf[:, q] -= fk
    assert np.max(np.abs(ring_sum(orc, m, pos, P, n, strict) - full)) > 1e-3 * np.max(np.abs(full))
