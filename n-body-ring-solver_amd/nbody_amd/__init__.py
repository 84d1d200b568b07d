"""Python view of the gfx950 N-body solver's C ABI (include/nbody.h).

A thin ctypes layer: every call goes to lib/libnbody_hip.so, which holds the
HIP kernels.  There is no Python or CPU fallback -- if the library is missing
or no gfx950 device is present, the calls raise.

Arrays follow the reference's struct-of-arrays layout (CPU/nbody_system.h:41-74):
positions, velocities, forces and accelerations are (3, N) float64 arrays and
masses are (N,).  Ring-share particle sets are (n, 4) rows of (x, y, z, m), the
FPGA particle_t layout (FPGA/host/solver.cpp:721).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

__all__ = [
    "LIB_PATH", "NBodyError", "Solver", "load_library", "init_system", "yoshida", "FORMS", "IC_KINDS",
    "snapshot_write", "snapshot_frames", "snapshot_read",
    "mlp_param_count",
]

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libnbody_hip.so")

FORMS = {"fp64": 0, "mixed": 1, "learned": 2, "mixed-raw": 3, "fp32": 4}
IC_KINDS = {"ring": 0, "sphere": 1, "plummer": 2}
STATUS = {0: "NB_OK", 1: "NB_ERR_ARG", 2: "NB_ERR_DEVICE", 3: "NB_ERR_ALLOC", 4: "NB_ERR_STATE",
          5: "NB_ERR_LAUNCH"}

# Every symbol include/nbody.h declares (checked by the CPU test suite).
EXPORTS = (
    "nb_last_error", "nb_version", "nb_create", "nb_destroy", "nb_init_system", "nb_yoshida",
    "nb_set_state", "nb_get_state", "nb_accel", "nb_compute_forces", "nb_accel_device",
    "nb_accel_block", "nb_accel_block_device", "nb_step", "nb_energy", "nb_timing_enable",
    "nb_timing_read", "nb_timing_reset", "nb_synchronize", "nb_learned_configure",
    "nb_learned_set_mode", "nb_learned_collect", "nb_learned_train", "nb_learned_stats_read",
    "nb_mlp_param_count", "nb_mlp_get_params", "nb_mlp_set_params", "nb_mlp_forward", "nb_mlp_grad",
    "nb_set_kernel", "nb_learned_reset", "nb_learned_set_flag_ratio", "nb_learned_set_near_radius",
    "nb_near_accel", "nb_near_accel_kept", "nb_learned_set_hold_ratio", "nb_learned_set_rough_share",
    "nb_snapshot_write", "nb_snapshot_frames", "nb_snapshot_read",
    "nb_step_euler", "nb_runge_kutta", "nb_step_rk", "nb_force_error", "nb_deviation", "nb_inertia",
    "nb_ring_predict", "nb_kernel_info_read", "nb_mlp_adam_step", "nb_forces_ring_round",
    "nb_forces_ring_round_device", "nb_accel_block_device_noalloc", "nb_run_records",
)
KERNELS = {"auto": 0, "tile": 1, "symmetric": 2}

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_P3 = C.POINTER(_D)
_SZ = C.c_size_t
_VP = C.c_void_p


class NBodyError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: {STATUS.get(status, status)}: {msg}")
        self.status = status


class LearnedStats(C.Structure):
    _fields_ = [("exact_evals", C.c_uint64), ("predicted_evals", C.c_uint64), ("fallbacks", C.c_uint64),
                ("subset_rows", C.c_uint64), ("last_check_err", C.c_double), ("max_check_err", C.c_double),
                ("net_evals", C.c_uint64), ("last_cubic_err", C.c_double), ("last_check_max", C.c_double),
                ("checks", C.c_uint64), ("net_checks", C.c_uint64), ("net_cubic_err_sum", C.c_double),
                ("net_err_sum", C.c_double), ("rough_exact", C.c_uint64)]


class KernelInfo(C.Structure):
    _fields_ = [("kernel", C.c_int), ("form", C.c_int), ("per_lane", C.c_int), ("block", C.c_int),
                ("items", C.c_uint64), ("scratch_bytes", C.c_uint64), ("name", C.c_char * 64)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load (once) and prototype libnbody_hip.so; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(path)
    L.nb_last_error.restype = C.c_char_p
    L.nb_version.restype = C.c_char_p
    L.nb_create.argtypes = [C.POINTER(_VP), C.c_int]
    L.nb_destroy.argtypes = [_VP]
    L.nb_init_system.argtypes = [C.c_int, _SZ, C.c_uint64, _D, _D, _P3, _P3, _P3]
    L.nb_yoshida.argtypes = [C.c_int, _D, _D, C.POINTER(_SZ)]
    L.nb_set_state.argtypes = [_VP, _SZ, _D, _P3, _P3]
    L.nb_get_state.argtypes = [_VP, _P3, _P3]
    L.nb_accel.argtypes = [_VP, C.c_int, C.c_double, _P3]
    L.nb_compute_forces.argtypes = [_VP, C.c_int, C.c_double, _P3]
    L.nb_accel_device.argtypes = [_VP, C.c_int, C.c_double]
    L.nb_accel_block.argtypes = [_VP, C.c_int, C.c_double, _SZ, _D, _SZ, _D, _SZ, _D]
    L.nb_accel_block_device.argtypes = [_VP, C.c_int, C.c_double, _SZ, _VP, _SZ, _VP, _SZ, _VP]
    L.nb_accel_block_device_noalloc.argtypes = [_VP, C.c_int, C.c_double, _SZ, _VP, _SZ, _VP, _SZ, _VP]
    L.nb_step.argtypes = [_VP, _SZ, C.c_double, _D, _D, _SZ, C.c_int, C.c_double]
    L.nb_energy.argtypes = [_VP, C.c_double, _D, _D]
    L.nb_timing_enable.argtypes = [_VP, C.c_int]
    L.nb_timing_read.argtypes = [_VP, _D, C.POINTER(C.c_uint64), _D]
    L.nb_timing_reset.argtypes = [_VP]
    L.nb_synchronize.argtypes = [_VP]
    L.nb_set_kernel.argtypes = [_VP, C.c_int]
    L.nb_learned_configure.argtypes = [_VP, C.c_int, C.c_double, C.c_int]
    L.nb_learned_set_mode.argtypes = [_VP, C.c_int]
    L.nb_learned_reset.argtypes = [_VP, C.c_uint64]
    L.nb_learned_set_flag_ratio.argtypes = [_VP, C.c_double]
    L.nb_learned_set_near_radius.argtypes = [_VP, C.c_double]
    L.nb_near_accel.argtypes = [_VP, C.c_int, C.c_double, C.c_double, _P3]
    L.nb_near_accel_kept.argtypes = [_VP, C.c_int, C.c_double, C.c_double, _P3]
    L.nb_learned_set_hold_ratio.argtypes = [_VP, C.c_double]
    L.nb_learned_set_rough_share.argtypes = [_VP, C.c_double]
    L.nb_snapshot_write.argtypes = [C.c_char_p, C.c_int, C.c_double, C.c_size_t, _D, _P3, _P3, _P3]
    L.nb_snapshot_frames.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
    L.nb_snapshot_read.argtypes = [C.c_char_p, C.c_size_t, C.c_size_t, _D, _D, _P3, _P3, _P3]
    L.nb_run_records.argtypes = [_VP, C.c_int, C.c_double, C.c_double, _D, _D, _D, _D, _D, _SZ, _SZ, C.c_double,
                                 _D, _D, _SZ, _D]
    L.nb_step_euler.argtypes = [_VP, _SZ, C.c_double, C.c_int, C.c_double]
    L.nb_runge_kutta.argtypes = [C.c_int, _D, _D, C.POINTER(_SZ)]
    L.nb_step_rk.argtypes = [_VP, _SZ, C.c_double, _D, _D, _SZ, C.c_int, C.c_double]
    L.nb_force_error.argtypes = [_SZ, _P3, _P3, _D]
    L.nb_deviation.argtypes = [_SZ, _P3, _P3, _D]
    L.nb_inertia.argtypes = [_SZ, _D, _P3, _D]
    L.nb_ring_predict.argtypes = [_SZ, C.c_double, C.c_double, _SZ, _SZ, _P3, _P3]
    L.nb_kernel_info_read.argtypes = [_VP, C.POINTER(KernelInfo)]
    L.nb_forces_ring_round.argtypes = [_VP, C.c_int, C.c_double, _SZ, _D, _P3, _P3, _P3, _P3, _SZ, _SZ, _SZ]
    L.nb_forces_ring_round_device.argtypes = [_VP, C.c_int, C.c_double, _SZ, _VP, _VP, _VP, _VP, _SZ, _SZ, _SZ]
    L.nb_learned_collect.argtypes = [_VP, _SZ, C.c_double, _D, _D, _SZ, C.c_int, C.c_double]
    L.nb_learned_train.argtypes = [_VP, C.c_int, C.c_int, C.c_double, C.c_uint64, _D]
    L.nb_learned_stats_read.argtypes = [_VP, C.POINTER(LearnedStats)]
    L.nb_mlp_param_count.argtypes = [C.POINTER(_SZ)]
    L.nb_mlp_get_params.argtypes = [_VP, _F]
    L.nb_mlp_set_params.argtypes = [_VP, _F]
    L.nb_mlp_forward.argtypes = [_VP, _SZ, _F, _F]
    L.nb_mlp_grad.argtypes = [_VP, _SZ, _F, _F, _F, _D]
    L.nb_mlp_adam_step.argtypes = [_VP, _F, C.c_double]
    for name in EXPORTS:
        if name not in ("nb_last_error", "nb_version"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _check(status: int, where: str):
    if status != 0:
        raise NBodyError(status, where, _lib.nb_last_error().decode())


def _p(a: np.ndarray, dtype=np.float64):
    assert a.dtype == dtype and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_D if dtype == np.float64 else _F)


def _planes(a: np.ndarray):
    """(3, N) C-contiguous float64 -> double*[3]."""
    assert a.ndim == 2 and a.shape[0] == 3 and a.dtype == np.float64 and a.flags.c_contiguous
    return (_D * 3)(*[a[k].ctypes.data_as(_D) for k in range(3)])


def _form(form) -> int:
    if isinstance(form, str):
        return FORMS[form]
    return int(form)


def snapshot_write(path: str, mass, pos, vel, force=None, t: float = 0.0, append: bool = False):
    """Append (or start) a frame in the reference's NBody_<N>.dat layout."""
    L = load_library()
    pos, vel = np.ascontiguousarray(pos, dtype=np.float64), np.ascontiguousarray(vel, dtype=np.float64)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    f = None if force is None else np.ascontiguousarray(force, dtype=np.float64)
    _check(L.nb_snapshot_write(os.fsencode(path), int(append), t, mass.shape[0], _p(mass), _planes(pos), _planes(vel),
                               None if f is None else _planes(f)), "nb_snapshot_write")


def snapshot_frames(path: str, n: int) -> int:
    L = load_library()
    k = C.c_size_t()
    _check(L.nb_snapshot_frames(os.fsencode(path), n, C.byref(k)), "nb_snapshot_frames")
    return k.value


def snapshot_read(path: str, n: int, frame: int = 0):
    """Frame `frame` of a snapshot file: (t, mass (N,), pos, vel, force (3, N))."""
    L = load_library()
    t = C.c_double()
    m = np.empty(n)
    pos, vel, force = np.empty((3, n)), np.empty((3, n)), np.empty((3, n))
    _check(L.nb_snapshot_read(os.fsencode(path), n, frame, C.byref(t), _p(m), _planes(pos), _planes(vel),
                              _planes(force)), "nb_snapshot_read")
    return t.value, m, pos, vel, force


def init_system(kind: str, n: int, seed: int = 42, params=None):
    """Seeded initial conditions: returns mass (N,), pos (3,N), vel (3,N), force (3,N) or None."""
    L = load_library()
    m = np.empty(n)
    pos, vel = np.empty((3, n)), np.empty((3, n))
    force = np.empty((3, n)) if kind == "ring" else None
    par = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    _check(L.nb_init_system(IC_KINDS[kind], n, seed, None if par is None else _p(par), _p(m), _planes(pos),
                            _planes(vel), None if force is None else _planes(force)), "nb_init_system")
    return m, pos, vel, force


def yoshida(order: int):
    L = load_library()
    c, d = np.zeros(16), np.zeros(16)
    lf = _SZ()
    _check(L.nb_yoshida(order, _p(c), _p(d), C.byref(lf)), "nb_yoshida")
    return c[:lf.value].copy(), d[:lf.value - 1].copy(), lf.value


def runge_kutta(order: int = 4):
    """The reference's Runge-Kutta tableau: (a ((s-1)^2,), b (s,), s)."""
    L = load_library()
    a, b = np.zeros(49), np.zeros(8)
    s = _SZ()
    _check(L.nb_runge_kutta(order, _p(a), _p(b), C.byref(s)), "nb_runge_kutta")
    return a[:(s.value - 1) ** 2].copy(), b[:s.value].copy(), s.value


def force_error(calc, ref) -> np.ndarray:
    """The reference's six force-error metrics of calc against ref ((3, N) each)."""
    L = load_library()
    calc = np.ascontiguousarray(calc, dtype=np.float64)
    ref = np.ascontiguousarray(ref, dtype=np.float64)
    out = np.empty(6)
    _check(L.nb_force_error(calc.shape[1], _planes(calc), _planes(ref), _p(out)), "nb_force_error")
    return out


def deviation(calc, ref) -> np.ndarray:
    """(mean |calc - ref|, max relative length deviation)."""
    L = load_library()
    calc = np.ascontiguousarray(calc, dtype=np.float64)
    ref = np.ascontiguousarray(ref, dtype=np.float64)
    out = np.empty(2)
    _check(L.nb_deviation(calc.shape[1], _planes(calc), _planes(ref), _p(out)), "nb_deviation")
    return out


def inertia(mass, pos) -> float:
    L = load_library()
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    out = C.c_double()
    _check(L.nb_inertia(mass.shape[0], _p(mass), _planes(pos), C.byref(out)), "nb_inertia")
    return out.value


def ring_predict(n: int, R: float, omega: float, step: int, steps_per_rev: int):
    """Rigid-rotation positions and velocities of the ring after step/steps_per_rev revolutions."""
    L = load_library()
    pos, vel = np.empty((3, n)), np.empty((3, n))
    _check(L.nb_ring_predict(n, R, omega, step, steps_per_rev, _planes(pos), _planes(vel)), "nb_ring_predict")
    return pos, vel


def mlp_param_count() -> int:
    L = load_library()
    n = _SZ()
    _check(L.nb_mlp_param_count(C.byref(n)), "nb_mlp_param_count")
    return n.value


class Solver:
    """One context on one GPU: the device-resident system and its stream."""

    def __init__(self, device: int = 0):
        self.L = load_library()
        h = _VP()
        _check(self.L.nb_create(C.byref(h), device), "nb_create")
        self.h = h
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.nb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- state ---------------------------------------------------------------
    def set_state(self, mass, pos, vel=None):
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        n = mass.shape[0]
        assert pos.shape == (3, n)
        vp = None
        if vel is not None:
            vel = np.ascontiguousarray(vel, dtype=np.float64)
            assert vel.shape == (3, n)
            vp = _planes(vel)
        _check(self.L.nb_set_state(self.h, n, _p(mass), _planes(pos), vp), "nb_set_state")
        self.n = n

    def get_state(self):
        pos, vel = np.empty((3, self.n)), np.empty((3, self.n))
        _check(self.L.nb_get_state(self.h, _planes(pos), _planes(vel)), "nb_get_state")
        return pos, vel

    # --- forces --------------------------------------------------------------
    def accel(self, form="fp64", eps: float = 0.0) -> np.ndarray:
        a = np.empty((3, self.n))
        _check(self.L.nb_accel(self.h, _form(form), eps, _planes(a)), "nb_accel")
        return a

    def compute_forces(self, form="fp64", eps: float = 0.0) -> np.ndarray:
        f = np.empty((3, self.n))
        _check(self.L.nb_compute_forces(self.h, _form(form), eps, _planes(f)), "nb_compute_forces")
        return f

    def accel_device(self, form="fp64", eps: float = 0.0):
        _check(self.L.nb_accel_device(self.h, _form(form), eps), "nb_accel_device")

    def accel_block(self, pi, pj, form="fp64", eps: float = 0.0, j_block: int = 0) -> np.ndarray:
        pi = np.ascontiguousarray(pi, dtype=np.float64)
        pj = np.ascontiguousarray(pj, dtype=np.float64)
        assert pi.ndim == 2 and pi.shape[1] == 4 and pj.ndim == 2 and pj.shape[1] == 4
        out = np.empty((pi.shape[0], 3))
        _check(self.L.nb_accel_block(self.h, _form(form), eps, pi.shape[0], _p(pi), pj.shape[0], _p(pj), j_block,
                                     _p(out)), "nb_accel_block")
        return out

    def accel_block_device(self, form, eps, ni, d_pi, nj, d_pj, j_block, d_out):
        """Device pointers (ints, e.g. torch tensor.data_ptr()); queued, not synchronised."""
        _check(self.L.nb_accel_block_device(self.h, _form(form), eps, ni, d_pi, nj, d_pj, j_block, d_out),
               "nb_accel_block_device")

    def forces_ring_round(self, mass, rank: int, comm_sz: int, recv_round: int, local_pos, recv_pos, force,
                          recv_force, form="fp64", eps: float = 0.0):
        """One ring round in the reference's reduced form; force and recv_force
        ((3, local_n) float64, C-contiguous) accumulate in place."""
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        lp = np.ascontiguousarray(local_pos, dtype=np.float64)
        rp = np.ascontiguousarray(recv_pos, dtype=np.float64)
        n = lp.shape[1]
        assert rp.shape == lp.shape == force.shape == recv_force.shape and mass.shape == (comm_sz * n,)
        _check(self.L.nb_forces_ring_round(self.h, _form(form), eps, n, _p(mass), _planes(lp), _planes(force),
                                           _planes(rp), _planes(recv_force), rank, comm_sz, recv_round),
               "nb_forces_ring_round")

    def forces_ring_round_device(self, form, eps, n, d_local, d_local_force4, d_recv, d_recv_force4, rank, comm_sz,
                                 recv_round):
        """Device pointers (ints); queued, not synchronised."""
        _check(self.L.nb_forces_ring_round_device(self.h, _form(form), eps, n, d_local, d_local_force4, d_recv,
                                                  d_recv_force4, rank, comm_sz, recv_round),
               "nb_forces_ring_round_device")

    def accel_block_device_noalloc(self, form, eps, ni, d_pi, nj, d_pj, j_block, d_out):
        """As accel_block_device with the partial sums in the library's static
        device scratch (nothing allocated); queued, not synchronised."""
        _check(self.L.nb_accel_block_device_noalloc(self.h, _form(form), eps, ni, d_pi, nj, d_pj, j_block, d_out),
               "nb_accel_block_device_noalloc")

    # --- integration ---------------------------------------------------------
    def step(self, steps: int, dt: float, order: int = 2, form="fp64", eps: float = 0.0, coeffs=None):
        c, d, lf = coeffs if coeffs is not None else yoshida(order)
        c = np.ascontiguousarray(c, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        _check(self.L.nb_step(self.h, steps, dt, _p(c), _p(d), lf, _form(form), eps), "nb_step")

    def run_records(self, mass, pos4, vel4, steps: int, dt: float, order: int = 2, form="fp64",
                    eps: float = 0.0, K: float = 1.0, coeff=None, force4=None, coeffs=None) -> float:
        """Whole run in the FPGA host's record form (nb_run_records): pos4 and
        vel4 ((n, 4) float64, C-contiguous; column 3 is neither read nor
        written) advance in place, force4 (same shape, optional) receives
        K coeff_i a_i of the last force evaluation; returns the device
        seconds of the steps."""
        for a in (pos4, vel4) + (() if force4 is None else (force4,)):
            assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"] and a.ndim == 2 and a.shape[1] == 4
        n = pos4.shape[0]
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        cf = None if coeff is None else np.ascontiguousarray(coeff, dtype=np.float64)
        assert vel4.shape == pos4.shape and mass.shape == (n,) and (cf is None or cf.shape == (n,))
        c, d, lf = coeffs if coeffs is not None else yoshida(order)
        c = np.ascontiguousarray(c, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        secs = C.c_double()
        _check(self.L.nb_run_records(self.h, _form(form), eps, K, _p(mass), None if cf is None else _p(cf),
                                     _p(pos4), _p(vel4), None if force4 is None else _p(force4), n, steps, dt,
                                     _p(c), _p(d), lf, C.byref(secs)), "nb_run_records")
        return secs.value

    def step_euler(self, steps: int, dt: float, form="fp64", eps: float = 0.0):
        _check(self.L.nb_step_euler(self.h, steps, dt, _form(form), eps), "nb_step_euler")

    def step_rk(self, steps: int, dt: float, form="fp64", eps: float = 0.0, tableau=None):
        a, b, s = tableau if tableau is not None else runge_kutta(4)
        a = np.ascontiguousarray(a if len(a) else np.zeros(1), dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        _check(self.L.nb_step_rk(self.h, steps, dt, _p(a), _p(b), s, _form(form), eps), "nb_step_rk")

    def energy(self, eps: float = 0.0):
        T, U = C.c_double(), C.c_double()
        _check(self.L.nb_energy(self.h, eps, C.byref(T), C.byref(U)), "nb_energy")
        return T.value, U.value

    # --- timing --------------------------------------------------------------
    def timing(self, enable: bool = True):
        _check(self.L.nb_timing_enable(self.h, int(enable)), "nb_timing_enable")

    def timing_read(self):
        ms, st = C.c_double(), C.c_double()
        ev = C.c_uint64()
        _check(self.L.nb_timing_read(self.h, C.byref(ms), C.byref(ev), C.byref(st)), "nb_timing_read")
        return ms.value, ev.value, st.value

    def timing_reset(self):
        _check(self.L.nb_timing_reset(self.h), "nb_timing_reset")

    def set_kernel(self, kernel: str):
        """Force kernel for nb_accel / nb_step: "auto", "tile" or "symmetric"."""
        _check(self.L.nb_set_kernel(self.h, KERNELS[kernel]), "nb_set_kernel")

    def synchronize(self):
        _check(self.L.nb_synchronize(self.h), "nb_synchronize")

    def kernel_info(self) -> dict:
        """What the last force evaluation of the resident system ran (kernel, form, name, scratch bytes)."""
        k = KernelInfo()
        _check(self.L.nb_kernel_info_read(self.h, C.byref(k)), "nb_kernel_info_read")
        d = {f: getattr(k, f) for f, _ in KernelInfo._fields_}
        d["name"] = k.name.decode()
        d["kernel"] = {v: n for n, v in KERNELS.items()}.get(k.kernel, k.kernel)
        d["form"] = {v: n for n, v in FORMS.items()}.get(k.form, k.form)
        return d

    # --- learned extrapolator -----------------------------------------------
    def learned_configure(self, period: int = 4, threshold: float = 1e-3, exact_form="mixed"):
        _check(self.L.nb_learned_configure(self.h, period, threshold, _form(exact_form)), "nb_learned_configure")

    def learned_reset(self, seed: int = 1):
        _check(self.L.nb_learned_reset(self.h, seed), "nb_learned_reset")

    def learned_set_flag_ratio(self, rho: float):
        _check(self.L.nb_learned_set_flag_ratio(self.h, rho), "nb_learned_set_flag_ratio")

    def learned_set_near_radius(self, rc: float):
        _check(self.L.nb_learned_set_near_radius(self.h, rc), "nb_learned_set_near_radius")

    def learned_set_hold_ratio(self, kappa: float):
        _check(self.L.nb_learned_set_hold_ratio(self.h, kappa), "nb_learned_set_hold_ratio")

    def learned_set_rough_share(self, share: float):
        _check(self.L.nb_learned_set_rough_share(self.h, share), "nb_learned_set_rough_share")

    def near_accel(self, rc: float, eps: float = 0.0, form="fp64") -> np.ndarray:
        a = np.empty((3, self.n))
        _check(self.L.nb_near_accel(self.h, _form(form), rc, eps, _planes(a)), "nb_near_accel")
        return a

    def near_accel_kept(self, rc: float, eps: float = 0.0, form="fp64") -> np.ndarray:
        """The near field on the grid of the last build (same n and rc), searched
        wider by the displacement since."""
        a = np.empty((3, self.n))
        _check(self.L.nb_near_accel_kept(self.h, _form(form), rc, eps, _planes(a)), "nb_near_accel_kept")
        return a

    def learned_set_mode(self, mode: int):
        _check(self.L.nb_learned_set_mode(self.h, mode), "nb_learned_set_mode")

    def learned_collect(self, steps: int, dt: float, order: int = 2, form="mixed", eps: float = 0.0):
        c, d, lf = yoshida(order)
        _check(self.L.nb_learned_collect(self.h, steps, dt, _p(c), _p(d), lf, _form(form), eps),
               "nb_learned_collect")

    def learned_train(self, iters: int, batch: int = 4096, lr: float = 1e-3, seed: int = 1) -> float:
        loss = C.c_double()
        _check(self.L.nb_learned_train(self.h, iters, batch, lr, seed, C.byref(loss)), "nb_learned_train")
        return loss.value

    def learned_stats(self) -> dict:
        s = LearnedStats()
        _check(self.L.nb_learned_stats_read(self.h, C.byref(s)), "nb_learned_stats_read")
        return {k: getattr(s, k) for k, _ in LearnedStats._fields_}

    def mlp_get_params(self) -> np.ndarray:
        out = np.empty(mlp_param_count(), dtype=np.float32)
        _check(self.L.nb_mlp_get_params(self.h, _p(out, np.float32)), "nb_mlp_get_params")
        return out

    def mlp_set_params(self, params: np.ndarray):
        params = np.ascontiguousarray(params, dtype=np.float32)
        assert params.shape == (mlp_param_count(),)
        _check(self.L.nb_mlp_set_params(self.h, _p(params, np.float32)), "nb_mlp_set_params")

    def mlp_forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        assert x.ndim == 2 and x.shape[1] == 15
        y = np.empty((x.shape[0], 3), dtype=np.float32)
        _check(self.L.nb_mlp_forward(self.h, x.shape[0], _p(x, np.float32), _p(y, np.float32)), "nb_mlp_forward")
        return y

    def mlp_grad(self, x: np.ndarray, target: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        g = np.empty(mlp_param_count(), dtype=np.float32)
        loss = C.c_double()
        _check(self.L.nb_mlp_grad(self.h, x.shape[0], _p(x, np.float32), _p(target, np.float32), _p(g, np.float32),
                                  C.byref(loss)), "nb_mlp_grad")
        return g, loss.value

    def mlp_adam_step(self, grads: np.ndarray, lr: float):
        """One Adam step of the network's parameters with a host gradient (test hook)."""
        grads = np.ascontiguousarray(grads, dtype=np.float32)
        assert grads.shape == (mlp_param_count(),)
        _check(self.L.nb_mlp_adam_step(self.h, _p(grads, np.float32), lr), "nb_mlp_adam_step")
