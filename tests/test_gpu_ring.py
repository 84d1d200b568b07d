"""One ring round in the reference's reduced form on the GPU
(nb_forces_ring_round, replacing compute_forces_avx2_blocked /
compute_forces_avx512_rsqrt_4i_blocked as hybrid_ring_solver calls them per
received block, CPU/hybrid_solver.cpp:84-95): against the reference's own
outputs for every (rank, recv_round) of 2- and 3-rank rings, against the
oracle, and at full size through the ring invariant (all rounds at all
ranks sum to the whole system's force)."""
import numpy as np
import pytest

import nbody_amd
from conftest import golden

pytestmark = pytest.mark.gpu


def increments_err(out, ref, base):
    """max over particles of |out - ref| / |ref - base| (the part the round added)."""
    d = np.linalg.norm(out - ref, axis=0)
    s = np.linalg.norm(ref - base, axis=0)
    mask = s > 0
    assert np.all(d[~mask] <= 1e-300)
    return float(np.max(d[mask] / s[mask])) if mask.any() else 0.0


@pytest.mark.parametrize("name", ["ringround_64_p2.npz", "ringround_300_p3.npz"])
@pytest.mark.parametrize("variant", ["avx2", "avx512"])
def test_ring_round_vs_reference(solver, name, variant):
    g = golden(name)
    P, n = g["force0"].shape[0], g["force0"].shape[3]
    m, pos = g["mass"], g["pos"]
    blocks = [np.ascontiguousarray(pos[:, r * n:(r + 1) * n]) for r in range(P)]
    worst = 0.0
    for rank in range(P):
        for rr in range(P):
            f = np.ascontiguousarray(g["force0"][rank, rr].copy())
            r = np.ascontiguousarray(g["recv0"][rank, rr].copy())
            solver.forces_ring_round(m, rank, P, rr, blocks[rank], blocks[(rank + rr) % P], f, r, "fp64", 0.0)
            worst = max(worst, increments_err(f, g[f"force_{variant}"][rank, rr], g["force0"][rank, rr]),
                        increments_err(r, g[f"recv_{variant}"][rank, rr], g["recv0"][rank, rr]))
    assert worst < 1e-12


@pytest.mark.parametrize("eps", [0.0, 0.01])
@pytest.mark.parametrize("form,tol", [("fp64", 1e-12), ("mixed", 2e-5)])
def test_ring_round_vs_oracle(solver, orc, eps, form, tol):
    rng = np.random.default_rng(3)
    P, n = 3, 1000
    m = rng.uniform(0.5, 2.0, P * n) / (P * n)
    _, pos, _, _ = nbody_amd.init_system("plummer", P * n, 9, [1.0, 20.0])
    blocks = [np.ascontiguousarray(pos[:, r * n:(r + 1) * n]) for r in range(P)]
    worst = 0.0
    for rank in range(P):
        for rr in range(P):
            f0, r0 = rng.standard_normal((3, n)) * 1e-3, rng.standard_normal((3, n)) * 1e-3
            fg, rg, fo, ro = f0.copy(), r0.copy(), f0.copy(), r0.copy()
            other = blocks[(rank + rr) % P]
            solver.forces_ring_round(m, rank, P, rr, blocks[rank], other, fg, rg, form, eps)
            orc.ring_round(m, rank, P, rr, blocks[rank], other, fo, ro, eps)
            worst = max(worst, increments_err(fg, fo, f0), increments_err(rg, ro, r0))
    assert worst < tol


@pytest.mark.parametrize("n,equal", [(131072, True), (131072 + 77, False)])
@pytest.mark.parametrize("form,tol", [("fp64", 1e-12), ("mixed", 2e-5)])
def test_ring_invariant_full_size(solver, orc, form, tol, n, equal):
    """Two ranks of 131072 (the ring share's block size) over a Plummer
    N = 262144: the local forces of both rounds plus the received forces
    that come back equal the whole system's forces (nb_compute_forces).
    Also two ranks of 131149 (ragged blocks) with masses apart, whose summed
    forces are checked against the oracle on sampled rows as well."""
    P = 2
    m, pos, _, _ = nbody_amd.init_system("plummer", P * n, 5, [1.0, 20.0])
    if not equal:
        m = np.random.default_rng(n).uniform(0.5, 1.5, P * n) / (P * n)
    blocks = [np.ascontiguousarray(pos[:, r * n:(r + 1) * n]) for r in range(P)]
    F = np.zeros((3, P * n))
    for rank in range(P):
        f = np.zeros((3, n))
        for rr in range(P):
            partner = (rank + rr) % P
            rf = np.zeros((3, n))
            solver.forces_ring_round(m, rank, P, rr, blocks[rank], blocks[partner], f, rf, form, 0.005)
            F[:, partner * n:(partner + 1) * n] += rf
        F[:, rank * n:(rank + 1) * n] += f
    solver.set_state(m, pos)
    full = solver.compute_forces("fp64", 0.005)
    err = np.max(np.linalg.norm(F - full, axis=0) / np.linalg.norm(full, axis=0))
    assert err < tol, err
    if not equal:
        rows = np.random.default_rng(2).choice(P * n, 128, replace=False)
        rows[:4] = [0, n - 1, n, P * n - 1]   # the first and the ragged last block of each rank
        pj = np.stack([pos[0], pos[1], pos[2], m], axis=1)
        ref = m[rows] * orc.accel_block(pj[rows], pj, 0.005).T
        err = np.max(np.linalg.norm(F[:, rows] - ref, axis=0) / np.linalg.norm(ref, axis=0))
        assert err < tol, err


def test_ring_round_device_matches_host(solver):
    from conftest import DeviceArray
    rng = np.random.default_rng(5)
    P, n = 2, 700
    m = rng.uniform(0.5, 2.0, P * n)
    pos = rng.standard_normal((3, P * n))
    lp, rp = np.ascontiguousarray(pos[:, :n]), np.ascontiguousarray(pos[:, n:])
    f0, r0 = rng.standard_normal((3, n)), rng.standard_normal((3, n))
    fh, rh = f0.copy(), r0.copy()
    solver.forces_ring_round(m, 0, P, 1, lp, rp, fh, rh, "fp64", 0.0)
    rec = lambda p, mm: DeviceArray(np.stack([p[0], p[1], p[2], mm], axis=1))
    bufs = dl, dr, fl, fr = rec(lp, m[:n]), rec(rp, m[n:]), rec(f0, np.zeros(n)), rec(r0, np.zeros(n))
    try:
        solver.forces_ring_round_device("fp64", 0.0, n, dl.addr, fl.addr, dr.addr, fr.addr, 0, P, 1)
        solver.synchronize()
        assert np.array_equal(fl.get()[:, :3].T, fh) and np.array_equal(fr.get()[:, :3].T, rh)
    finally:
        for b in bufs:
            b.free()


def test_ring_round_argument_errors(solver):
    z = np.zeros((3, 4))
    with pytest.raises(nbody_amd.NBodyError):
        solver.forces_ring_round(np.ones(8), 2, 2, 0, z, z, z.copy(), z.copy())   # rank >= comm_sz
    with pytest.raises(nbody_amd.NBodyError):
        solver.forces_ring_round(np.ones(8), 0, 2, 0, z, z, z.copy(), z.copy(), form="learned")
