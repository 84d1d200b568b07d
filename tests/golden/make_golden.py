"""Write the golden fixtures from the reference's own CPU routines.

Runs the reference's force kernels, leapfrog driver and energy functions
(compiled from its sources by oracle/Makefile into oracle/_ref/libnbref.so)
on small seeded inputs, and stores inputs and outputs as .npz files next to
this script.  Re-run with `python tests/golden/make_golden.py` after
`make -C oracle ref`; the fixtures are data (inputs and expected outputs)
and contain no reference code.

The reference has no softening, so every fixture is at eps = 0.

Fixtures
  ring_<N>.npz      reference ring + centre IC (CPU/nbody_system.cpp:235):
                    m, pos, vel, analytic force; forces of
                    compute_forces_avx2_blocked (force_avx2),
                    compute_forces_avx512_rsqrt_4i_blocked (force_avx512),
                    compute_forces_avx2_blocked_sp (force_sp)
  jitter_<N>.npz    the ring with seeded 3-D jitter (so no symmetry hides a
                    wrong sign or index), same forces, energies
                    (T, U_avx2, U_scalar), and 10 leapfrog steps of
                    reduced_solver_omp_lf at orders 2, 4, 6, 8
                    (pos_lf{k}, vel_lf{k}) with dt = 1e-3 and the Yoshida
                    coefficients the reference builds (c{k}, d{k}).
  drivers_<N>.npz   the same jittered system and 10 steps (dt = 1e-3) of the
                    reference's Euler driver reduced_solver_omp (pos_euler,
                    vel_euler) and of its Runge-Kutta driver
                    reduced_solver_omp_rk with the RK4 tableau of omp_main
                    (rk_a, rk_b; pos_rk4, vel_rk4).
  accuracy_<N>.npz  one omp-accuracy case: the ring at mass ratio 1e3 under a
                    fixed permutation (perm), its analytic force (force_ref)
                    and compute_forces_avx2_blocked's (force_avx2), the six
                    metrics of calculate_force_error on them (force_error),
                    calculate_deviation of jittered positions (pos_jit) from
                    the ring (deviation), compute_inertia (inertia), and the
                    rigid-rotation prediction after 37 of 1000 steps per
                    revolution (pos_pred, vel_pred).
  ringround_<n>_p<P>.npz
                    one ring round of the reduced scheme per (rank,
                    recv_round) of a P-rank ring over the jittered system of
                    P*n particles in contiguous blocks of n (mass, pos): the
                    force arrays the caller passes in (force0, recv0, seeded
                    non-zero so the accumulation shows) and what
                    compute_forces_avx2_blocked (force_avx2, recv_avx2) and
                    compute_forces_avx512_rsqrt_4i_blocked (force_avx512,
                    recv_avx512) leave in them; indexed [rank, recv_round].

Usage: python tests/golden/make_golden.py [group ...] with groups ring,
jitter, drivers, accuracy, ringround (default: all).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))

from oracle import Oracle, Reference  # noqa: E402


RK4_A = np.array([0.5, 0.0, 0.0, 0.0, 0.5, 0.0, 0.0, 0.0, 1.0])   # CPU/omp_solver.cpp:776-780
RK4_B = np.array([1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0])       # :781


def jittered(ref, n):
    m, pos, vel, _ = ref.init_ring(n, 1.0, 2 * np.pi, 1e3)
    rng = np.random.default_rng(1234 + n)
    pos = np.ascontiguousarray(pos + 0.05 * rng.standard_normal(pos.shape))
    vel = np.ascontiguousarray(vel + 0.05 * rng.standard_normal(vel.shape))
    # masses spread over two decades so a swapped m_i / m_j shows up
    m = np.ascontiguousarray(m * np.exp(rng.uniform(-2.3, 2.3, n)))
    return m, pos, vel


def write_drivers(ref):
    for n in (64, 300):
        m, pos, vel = jittered(ref, n)
        pe, ve = ref.euler(m, pos, vel, 10, 1e-3)
        pr, vr = ref.rk(m, pos, vel, 10, 1e-3, RK4_A, RK4_B)
        np.savez_compressed(os.path.join(HERE, f"drivers_{n}.npz"), m=m, pos=pos, vel=vel, pos_euler=pe,
                            vel_euler=ve, rk_a=RK4_A, rk_b=RK4_B, pos_rk4=pr, vel_rk4=vr)


def write_accuracy(ref):
    n = 257
    m, pos, vel, f = ref.init_ring(n, 1.0, 2 * np.pi, 1e3)
    perm = np.random.default_rng(77).permutation(n)
    m, pos, vel, f = m[perm].copy(), np.ascontiguousarray(pos[:, perm]), np.ascontiguousarray(vel[:, perm]), \
        np.ascontiguousarray(f[:, perm])
    fk = ref.forces(Reference.AVX2_BLOCKED, m, pos)
    pos_jit = np.ascontiguousarray(pos * (1 + 1e-4 * np.random.default_rng(78).standard_normal(pos.shape)))
    pp, vp = ref.ring_predict(n, 1.0, 2 * np.pi, 37, 1000)
    np.savez_compressed(os.path.join(HERE, f"accuracy_{n}.npz"), m=m, pos=pos, vel=vel, perm=perm, force_ref=f,
                        force_avx2=fk, force_error=ref.force_error(fk, f), pos_jit=pos_jit,
                        deviation=ref.deviation(pos_jit, pos), inertia=np.array(ref.inertia(m, pos_jit)),
                        pos_pred=pp, vel_pred=vp)


def write_ringround(ref):
    for n, P in ((64, 2), (300, 3)):
        m, pos, _ = jittered(ref, P * n)
        blocks = [np.ascontiguousarray(pos[:, r * n:(r + 1) * n]) for r in range(P)]
        rng = np.random.default_rng(4321 + n)
        out = {"mass": m, "pos": pos}
        shape = (P, P, 3, n)
        out["force0"] = rng.standard_normal(shape) * 1e-2
        out["recv0"] = rng.standard_normal(shape) * 1e-2
        for name, variant in (("avx2", Reference.AVX2_BLOCKED), ("avx512", Reference.AVX512_RSQRT_4I)):
            fo, ro = np.empty(shape), np.empty(shape)
            for rank in range(P):
                for rr in range(P):
                    f = out["force0"][rank, rr].copy()   # the round accumulates into its arrays
                    r = out["recv0"][rank, rr].copy()
                    ref.ring_round(variant, m, rank, P, rr, blocks[rank], blocks[(rank + rr) % P], f, r)
                    fo[rank, rr], ro[rank, rr] = f, r
            out[f"force_{name}"], out[f"recv_{name}"] = fo, ro
        np.savez_compressed(os.path.join(HERE, f"ringround_{n}_p{P}.npz"), **out)


def main(groups=("ring", "jitter", "drivers", "accuracy", "ringround")):
    ref = Reference()
    orc = Oracle()
    if "ringround" in groups:
        write_ringround(ref)
    if "drivers" in groups:
        write_drivers(ref)
    if "accuracy" in groups:
        write_accuracy(ref)
    for n in (64, 257, 2049) if "ring" in groups else ():
        m, pos, vel, f = ref.init_ring(n, 1.0, 2 * np.pi, 1e5)
        np.savez_compressed(
            os.path.join(HERE, f"ring_{n}.npz"), m=m, pos=pos, vel=vel, force_analytic=f,
            force_avx2=ref.forces(Reference.AVX2_BLOCKED, m, pos),
            force_avx512=ref.forces(Reference.AVX512_RSQRT_4I, m, pos),
            force_sp=ref.forces(Reference.AVX2_BLOCKED_SP, m, pos))
    for n in (64, 300) if "jitter" in groups else ():
        m, pos, vel = jittered(ref, n)
        out = dict(m=m, pos=pos, vel=vel,
                   force_avx2=ref.forces(Reference.AVX2_BLOCKED, m, pos),
                   force_avx512=ref.forces(Reference.AVX512_RSQRT_4I, m, pos),
                   force_sp=ref.forces(Reference.AVX2_BLOCKED_SP, m, pos))
        T, Ua, Us = ref.energy(m, pos, vel)
        out.update(T=T, U_avx2=Ua, U_scalar=Us)
        for order in (2, 4, 6, 8):
            c, d, lf = orc.yoshida(order)
            p1, v1 = ref.leapfrog(m, pos, vel, 10, 1e-3, c, d)
            out[f"pos_lf{order}"], out[f"vel_lf{order}"] = p1, v1
            out[f"c{order}"], out[f"d{order}"] = c, d
        np.savez_compressed(os.path.join(HERE, f"jitter_{n}.npz"), **out)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main(tuple(sys.argv[1:]) or ("ring", "jitter", "drivers", "accuracy", "ringround"))
