"""ctypes bindings for the CPU oracle (test infrastructure only).

Loads oracle/_build/liboracle.so (the C restatement) and, when present,
oracle/_ref/libnbref.so (the reference's own CPU sources compiled in place).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the solver itself never does.

Arrays follow the reference's struct-of-arrays layout (CPU/nbody_system.h:41-74):
positions, velocities and forces are (3, N) float64 arrays, masses (N,).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "_build", "liboracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libnbref.so")

_D = C.POINTER(C.c_double)
_SZ = C.c_size_t


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def _ptr3(a: np.ndarray):
    """(3, N) float64 -> double*[3] (the reference's `double* pos[DIM]`)."""
    assert a.shape[0] == 3 and a.dtype == np.float64 and a.flags.c_contiguous
    arr = (_D * 3)(*[a[k].ctypes.data_as(_D) for k in range(3)])
    return arr


class Oracle:
    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        self.lib = L = C.CDLL(path)
        L.orc_rng_stream.argtypes = [C.c_uint64, _SZ, _D]
        for name in ("orc_ic_sphere",):
            getattr(L, name).argtypes = [_SZ, C.c_uint64] + [_D] * 7
        L.orc_ic_plummer.argtypes = [_SZ, C.c_uint64, C.c_double, C.c_double] + [_D] * 7
        L.orc_ic_ring.argtypes = [_SZ, C.c_double, C.c_double, C.c_double] + [_D] * 10
        L.orc_ring_predict.argtypes = [_SZ, C.c_double, C.c_double, C.c_double] + [_D] * 6
        L.orc_ring_accel_exact.argtypes = [_SZ, C.c_double, C.c_double, C.c_double, C.c_double] + [_D] * 3
        for name in ("orc_accel", "orc_accel_ld", "orc_forces_reduced"):
            getattr(L, name).argtypes = [_SZ, _D, _D, _D, _D, C.c_double, C.c_double, _D, _D, _D]
        L.orc_accel_block.argtypes = [_SZ, _D, _SZ, _D, C.c_double, C.c_double, _D]
        L.orc_ring_round.argtypes = [_SZ, _D, _SZ, _SZ, _SZ] + [_D] * 6 + [C.c_double, C.c_double] + [_D] * 6
        L.orc_yoshida.argtypes = [C.c_int, _D, _D]
        L.orc_yoshida.restype = _SZ
        L.orc_leapfrog.argtypes = [_SZ, _D, _D, _D, _D, _D, _D, _D, _SZ, C.c_double, _D, _D, _SZ,
                                   C.c_double, C.c_double]
        L.orc_energy.argtypes = [_SZ, _D, _D, _D, _D, _D, _D, _D, C.c_double, C.c_double, _D, _D]
        L.orc_euler.argtypes = [_SZ, _D, _D, _D, _D, _D, _D, _D, _SZ, C.c_double, C.c_double, C.c_double]
        L.orc_rk.argtypes = [_SZ, _D, _D, _D, _D, _D, _D, _D, _SZ, C.c_double, _D, _D, _SZ, C.c_double,
                             C.c_double]
        L.orc_force_error.argtypes = [_SZ] + [_D] * 7
        L.orc_deviation.argtypes = [_SZ] + [_D] * 7
        L.orc_inertia.argtypes = [_SZ, _D, _D, _D, _D]
        L.orc_inertia.restype = C.c_double
        L.orc_max_rel_err.argtypes = [_SZ] + [_D] * 6 + [_D]
        L.orc_max_rel_err.restype = C.c_double
        L.orc_poly_extrap.argtypes = [_SZ, _D, _D, _D, _D, _D]
        L.orc_mlp_param_count.restype = _SZ
        L.orc_bf16_round.argtypes = [C.c_double]
        L.orc_bf16_round.restype = C.c_double
        L.orc_mlp_forward.argtypes = [_D, _SZ, _D, C.c_int, _D]
        L.orc_mlp_grad.argtypes = [_D, _SZ, _D, _D, C.c_int, _D]
        L.orc_mlp_grad.restype = C.c_double
        L.orc_adam.argtypes = [_SZ, _D, _D, _D, _D, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_features.argtypes = [_SZ, _D, _D, _D, _D, _D, _D, _D, _D, _D]
        L.orc_lagrange4.argtypes = [_D, C.c_double, _D]
        L.orc_near_accel.argtypes = [_SZ, _D, _D, _D, _D, C.c_double, C.c_double, _D, _D, _D]
        L.orc_near_taper.argtypes = [C.c_double, C.c_double]
        L.orc_near_taper.restype = C.c_double

    # --- initial conditions -------------------------------------------------
    def rng_stream(self, seed: int, n: int) -> np.ndarray:
        out = np.empty(n)
        self.lib.orc_rng_stream(seed, n, _p(out))
        return out

    def _alloc(self, n):
        return np.empty(n), np.empty((3, n)), np.empty((3, n))

    def ic_sphere(self, n: int, seed: int):
        m, pos, vel = self._alloc(n)
        self.lib.orc_ic_sphere(n, seed, _p(m), *[_p_row(pos, k) for k in range(3)],
                               *[_p_row(vel, k) for k in range(3)])
        return m, pos, vel

    def ic_plummer(self, n: int, seed: int, a: float = 1.0, rmax: float = 20.0):
        m, pos, vel = self._alloc(n)
        self.lib.orc_ic_plummer(n, seed, a, rmax, _p(m), *[_p_row(pos, k) for k in range(3)],
                                *[_p_row(vel, k) for k in range(3)])
        return m, pos, vel

    def ic_ring(self, n: int, R: float = 1.0, omega: float = 2 * np.pi, ratio: float = 1e5):
        m, pos, vel = self._alloc(n)
        f = np.empty((3, n))
        self.lib.orc_ic_ring(n, R, omega, ratio, _p(m), *[_p_row(pos, k) for k in range(3)],
                             *[_p_row(vel, k) for k in range(3)], *[_p_row(f, k) for k in range(3)])
        return m, pos, vel, f

    def ring_predict(self, n: int, R: float, omega: float, phase: float):
        pos, vel = np.empty((3, n)), np.empty((3, n))
        self.lib.orc_ring_predict(n, R, omega, phase, *[_p_row(pos, k) for k in range(3)],
                                  *[_p_row(vel, k) for k in range(3)])
        return pos, vel

    def ring_accel_exact(self, n: int, R: float, mring: float, mcentre: float, eps: float):
        a = np.empty((3, n))
        self.lib.orc_ring_accel_exact(n, R, mring, mcentre, eps, *[_p_row(a, k) for k in range(3)])
        return a

    # --- forces ---------------------------------------------------------------
    def accel(self, m, pos, eps: float, G: float = 1.0, long_double: bool = False) -> np.ndarray:
        n = m.shape[0]
        a = np.empty((3, n))
        fn = self.lib.orc_accel_ld if long_double else self.lib.orc_accel
        fn(n, _p(m), *[_p_row(pos, k) for k in range(3)], eps, G, *[_p_row(a, k) for k in range(3)])
        return a

    def forces_reduced(self, m, pos, eps: float, G: float = 1.0) -> np.ndarray:
        n = m.shape[0]
        f = np.empty((3, n))
        self.lib.orc_forces_reduced(n, _p(m), *[_p_row(pos, k) for k in range(3)], eps, G,
                                    *[_p_row(f, k) for k in range(3)])
        return f

    def ring_round(self, mass, rank: int, comm_sz: int, recv_round: int, local_pos, recv_pos, force,
                   recv_force, eps: float = 0.0, G: float = 1.0):
        """One ring round of the reduced scheme; force and recv_force ((3, n),
        float64, C-contiguous) accumulate in place."""
        n = local_pos.shape[1]
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        lp, rp = np.ascontiguousarray(local_pos), np.ascontiguousarray(recv_pos)
        assert mass.shape[0] == comm_sz * n and force.flags.c_contiguous and recv_force.flags.c_contiguous
        self.lib.orc_ring_round(n, _p(mass), rank, comm_sz, recv_round, *[_p_row(lp, k) for k in range(3)],
                                *[_p_row(rp, k) for k in range(3)], eps, G, *[_p_row(force, k) for k in range(3)],
                                *[_p_row(recv_force, k) for k in range(3)])

    def accel_block(self, pi: np.ndarray, pj: np.ndarray, eps: float, G: float = 1.0) -> np.ndarray:
        """pi (ni, 4), pj (nj, 4) rows of (x, y, z, m) -> (ni, 3) accelerations."""
        pi = np.ascontiguousarray(pi, dtype=np.float64)
        pj = np.ascontiguousarray(pj, dtype=np.float64)
        out = np.empty((pi.shape[0], 3))
        self.lib.orc_accel_block(pi.shape[0], _p(pi), pj.shape[0], _p(pj), eps, G, _p(out))
        return out

    # --- integrator ---------------------------------------------------------------
    def yoshida(self, order: int):
        c, d = np.zeros(16), np.zeros(16)
        lf = self.lib.orc_yoshida(order, _p(c), _p(d))
        return c[:lf].copy(), d[:lf - 1].copy(), lf

    def leapfrog(self, m, pos, vel, steps: int, dt: float, order: int = 2, eps: float = 0.0,
                 G: float = 1.0):
        c, d, lf = self.yoshida(order)
        pos = np.ascontiguousarray(pos.copy())
        vel = np.ascontiguousarray(vel.copy())
        self.lib.orc_leapfrog(m.shape[0], _p(m), *[_p_row(pos, k) for k in range(3)],
                              *[_p_row(vel, k) for k in range(3)], steps, dt, _p(c), _p(d), lf, eps, G)
        return pos, vel

    def euler(self, m, pos, vel, steps: int, dt: float, eps: float = 0.0, G: float = 1.0):
        pos = np.ascontiguousarray(pos.copy())
        vel = np.ascontiguousarray(vel.copy())
        self.lib.orc_euler(m.shape[0], _p(m), *[_p_row(pos, k) for k in range(3)],
                           *[_p_row(vel, k) for k in range(3)], steps, dt, eps, G)
        return pos, vel

    RK4_A = np.array([0.5, 0.0, 0.0, 0.0, 0.5, 0.0, 0.0, 0.0, 1.0])
    RK4_B = np.array([1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0])

    def rk(self, m, pos, vel, steps: int, dt: float, a=None, b=None, eps: float = 0.0, G: float = 1.0):
        """Explicit Runge-Kutta with tableau (a, b) in the reference's layout; classic RK4 by default."""
        a = np.ascontiguousarray(self.RK4_A if a is None else a, dtype=np.float64)
        b = np.ascontiguousarray(self.RK4_B if b is None else b, dtype=np.float64)
        s = b.shape[0]
        assert a.shape[0] >= (s - 1) ** 2
        pos = np.ascontiguousarray(pos.copy())
        vel = np.ascontiguousarray(vel.copy())
        self.lib.orc_rk(m.shape[0], _p(m), *[_p_row(pos, k) for k in range(3)],
                        *[_p_row(vel, k) for k in range(3)], steps, dt, _p(a if a.size else np.zeros(1)), _p(b),
                        s, eps, G)
        return pos, vel

    def force_error(self, calc, ref) -> np.ndarray:
        calc, ref = np.ascontiguousarray(calc), np.ascontiguousarray(ref)
        out = np.empty(6)
        self.lib.orc_force_error(calc.shape[1], *[_p_row(calc, k) for k in range(3)],
                                 *[_p_row(ref, k) for k in range(3)], _p(out))
        return out

    def deviation(self, calc, ref) -> np.ndarray:
        calc, ref = np.ascontiguousarray(calc), np.ascontiguousarray(ref)
        out = np.empty(2)
        self.lib.orc_deviation(calc.shape[1], *[_p_row(calc, k) for k in range(3)],
                               *[_p_row(ref, k) for k in range(3)], _p(out))
        return out

    def inertia(self, m, pos) -> float:
        pos = np.ascontiguousarray(pos)
        return self.lib.orc_inertia(m.shape[0], _p(m), *[_p_row(pos, k) for k in range(3)])

    def energy(self, m, pos, vel, eps: float, G: float = 1.0):
        T, U = C.c_double(), C.c_double()
        self.lib.orc_energy(m.shape[0], _p(m), *[_p_row(pos, k) for k in range(3)],
                            *[_p_row(vel, k) for k in range(3)], eps, G, C.byref(T), C.byref(U))
        return T.value, U.value

    def max_rel_err(self, a, ref):
        mean = C.c_double()
        mx = self.lib.orc_max_rel_err(a.shape[1], *[_p_row(a, k) for k in range(3)],
                                      *[_p_row(ref, k) for k in range(3)], C.byref(mean))
        return mx, mean.value

    # --- learned extrapolator --------------------------------------------------
    def mlp_param_count(self) -> int:
        return self.lib.orc_mlp_param_count()

    def bf16(self, v: float) -> float:
        return self.lib.orc_bf16_round(v)

    def mlp_forward(self, params, x, bf16: bool = True):
        """x (rows, 15) -> y (rows, 3); params packed W1 b1 W2 b2 W3 b3 W4 b4."""
        params = np.ascontiguousarray(params, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty((x.shape[0], 3))
        self.lib.orc_mlp_forward(_p(params), x.shape[0], _p(x), int(bf16), _p(y))
        return y

    def mlp_grad(self, params, x, t, bf16: bool = True):
        params = np.ascontiguousarray(params, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        t = np.ascontiguousarray(t, dtype=np.float64)
        g = np.empty_like(params)
        loss = self.lib.orc_mlp_grad(_p(params), x.shape[0], _p(x), _p(t), int(bf16), _p(g))
        return g, loss

    def adam(self, p, g, m, v, lr, t, b1=0.9, b2=0.999, eps=1e-8):
        for a in (p, g, m, v):
            assert a.dtype == np.float64 and a.flags.c_contiguous
        self.lib.orc_adam(p.shape[0], _p(p), _p(g), _p(m), _p(v), lr, b1, b2, eps, t)

    def features(self, a0, a1, a2, a3, v):
        """(n, 3) arrays -> features (n, 15), scale (n,), cubic (n, 3), frame (n, 3, 3)
        (frame[i, k] is the unit vector e_{k+1} of particle i)."""
        n = a0.shape[0]
        x, s, cub, fr = np.empty((n, 15)), np.empty(n), np.empty((n, 3)), np.empty((n, 3, 3))
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (a0, a1, a2, a3, v)]
        self.lib.orc_features(n, *[_p(a) for a in args], _p(x), _p(s), _p(cub), _p(fr))
        return x, s, cub, fr

    def lagrange4(self, t, x: float) -> np.ndarray:
        t = np.ascontiguousarray(t, dtype=np.float64)
        w = np.empty(4)
        self.lib.orc_lagrange4(_p(t), x, _p(w))
        return w

    def near_accel(self, m, pos, rc: float, eps: float) -> np.ndarray:
        n = m.shape[0]
        a = np.empty((3, n))
        self.lib.orc_near_accel(n, _p(m), *[_p_row(pos, k) for k in range(3)], rc, eps,
                                *[_p_row(a, k) for k in range(3)])
        return a

    def near_taper(self, r: float, rc: float) -> float:
        return self.lib.orc_near_taper(r, rc)

    def poly_extrap(self, a0, a1, a2, a3):
        out = np.empty_like(a0)
        self.lib.orc_poly_extrap(a0.size, *[_p(np.ascontiguousarray(v)) for v in (a0, a1, a2, a3)],
                                 _p(out))
        return out


def _p_row(a: np.ndarray, k: int):
    row = a[k]
    assert row.flags.c_contiguous and row.dtype == np.float64
    return row.ctypes.data_as(_D)


class Reference:
    """The reference's own CPU routines (compiled from its sources by oracle/Makefile)."""

    AVX2_BLOCKED, AVX512_RSQRT_4I, AVX512_RSQRT, AVX2, AVX2_BLOCKED_SP = range(5)

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref` with the reference present")
        self.lib = L = C.CDLL(path)
        P3 = C.POINTER(_D)
        L.ref_max_threads.restype = C.c_int
        L.ref_forces.argtypes = [C.c_int, _SZ, _D, P3, P3]
        L.ref_time_forces.argtypes = [C.c_int, _SZ, _D, P3, C.c_int]
        L.ref_time_forces.restype = C.c_double
        L.ref_leapfrog.argtypes = [_SZ, _D, P3, P3, _SZ, C.c_double, _D, _D, _SZ]
        L.ref_init_ring.argtypes = [_SZ, C.c_double, C.c_double, C.c_double, _D, P3, P3, P3]
        L.ref_energy.argtypes = [_SZ, _D, P3, P3, _D, _D, _D]
        L.ref_save_particles.argtypes = [_SZ, _D, P3, P3, P3, C.c_double, _SZ]
        L.ref_euler.argtypes = [_SZ, _D, P3, P3, _SZ, C.c_double]
        L.ref_rk.argtypes = [_SZ, _D, P3, P3, _SZ, C.c_double, _D, _D, _SZ]
        L.ref_force_error.argtypes = [_SZ, P3, P3, _D]
        L.ref_deviation.argtypes = [_SZ, P3, P3, _D]
        L.ref_inertia.argtypes = [_SZ, _D, P3]
        L.ref_inertia.restype = C.c_double
        L.ref_ring_predict.argtypes = [_SZ, C.c_double, C.c_double, _SZ, _SZ, P3, P3]
        L.ref_ring_round.argtypes = [C.c_int, _SZ, _SZ, _SZ, _SZ, _D, P3, P3, P3, P3]

    def forces(self, variant: int, m, pos) -> np.ndarray:
        n = m.shape[0]
        pos = np.ascontiguousarray(pos)
        f = np.empty((3, n))
        rc = self.lib.ref_forces(variant, n, _p(m), _ptr3(pos), _ptr3(f))
        assert rc == 0
        return f

    def ring_round(self, variant: int, mass, rank: int, comm_sz: int, recv_round: int, local_pos, recv_pos,
                   force, recv_force):
        """The reference kernel `variant` for one ring round; force and
        recv_force ((3, n), C-contiguous) accumulate in place."""
        lp, rp = np.ascontiguousarray(local_pos), np.ascontiguousarray(recv_pos)
        assert force.flags.c_contiguous and recv_force.flags.c_contiguous
        rc = self.lib.ref_ring_round(variant, lp.shape[1], comm_sz, rank, recv_round,
                                     _p(np.ascontiguousarray(mass, dtype=np.float64)), _ptr3(lp), _ptr3(rp),
                                     _ptr3(force), _ptr3(recv_force))
        assert rc == 0

    def time_forces(self, variant: int, m, pos, reps: int) -> float:
        pos = np.ascontiguousarray(pos)
        return self.lib.ref_time_forces(variant, m.shape[0], _p(m), _ptr3(pos), reps)

    def leapfrog(self, m, pos, vel, steps: int, dt: float, c, d):
        pos = np.ascontiguousarray(pos.copy())
        vel = np.ascontiguousarray(vel.copy())
        c = np.ascontiguousarray(c, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        rc = self.lib.ref_leapfrog(m.shape[0], _p(m), _ptr3(pos), _ptr3(vel), steps, dt, _p(c), _p(d),
                                   c.shape[0])
        assert rc == 0
        return pos, vel

    def save_particles(self, m, pos, vel, force, dt: float, frames: int):
        """NBody_system::save_particles for frames 0..frames-1 (velocities scaled by 1 + k) into
        NBody_<N>.dat in the current directory."""
        pos, vel, force = (np.ascontiguousarray(a) for a in (pos, vel, force))
        rc = self.lib.ref_save_particles(m.shape[0], _p(m), _ptr3(pos), _ptr3(vel), _ptr3(force), dt, frames)
        assert rc == 0

    def euler(self, m, pos, vel, steps: int, dt: float):
        pos = np.ascontiguousarray(pos.copy())
        vel = np.ascontiguousarray(vel.copy())
        assert self.lib.ref_euler(m.shape[0], _p(m), _ptr3(pos), _ptr3(vel), steps, dt) == 0
        return pos, vel

    def rk(self, m, pos, vel, steps: int, dt: float, a, b):
        pos = np.ascontiguousarray(pos.copy())
        vel = np.ascontiguousarray(vel.copy())
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        assert self.lib.ref_rk(m.shape[0], _p(m), _ptr3(pos), _ptr3(vel), steps, dt, _p(a), _p(b), b.shape[0]) == 0
        return pos, vel

    def force_error(self, calc, ref) -> np.ndarray:
        calc, ref = np.ascontiguousarray(calc), np.ascontiguousarray(ref)
        out = np.empty(6)
        self.lib.ref_force_error(calc.shape[1], _ptr3(calc), _ptr3(ref), _p(out))
        return out

    def deviation(self, calc, ref) -> np.ndarray:
        calc, ref = np.ascontiguousarray(calc), np.ascontiguousarray(ref)
        out = np.empty(2)
        self.lib.ref_deviation(calc.shape[1], _ptr3(calc), _ptr3(ref), _p(out))
        return out

    def inertia(self, m, pos) -> float:
        return self.lib.ref_inertia(m.shape[0], _p(m), _ptr3(np.ascontiguousarray(pos)))

    def ring_predict(self, n: int, R: float, omega: float, step: int, steps_per_rev: int):
        pos, vel = np.empty((3, n)), np.empty((3, n))
        self.lib.ref_ring_predict(n, R, omega, step, steps_per_rev, _ptr3(pos), _ptr3(vel))
        return pos, vel

    def init_ring(self, n: int, R: float = 1.0, omega: float = 2 * np.pi, ratio: float = 1e5):
        m = np.empty(n)
        pos, vel, f = np.empty((3, n)), np.empty((3, n)), np.empty((3, n))
        self.lib.ref_init_ring(n, R, omega, ratio, _p(m), _ptr3(pos), _ptr3(vel), _ptr3(f))
        return m, pos, vel, f

    def energy(self, m, pos, vel):
        T, Ua, Us = C.c_double(), C.c_double(), C.c_double()
        self.lib.ref_energy(m.shape[0], _p(m), _ptr3(np.ascontiguousarray(pos)),
                            _ptr3(np.ascontiguousarray(vel)), C.byref(T), C.byref(Ua), C.byref(Us))
        return T.value, Ua.value, Us.value


def cpu_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return " avx512f " in (line + " ")
    except OSError:
        pass
    return False
