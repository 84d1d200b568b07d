"""nb_accel_block_device_noalloc: the ring-share evaluation with its partial
sums in the library's static device scratch equals nb_accel_block_device's
result bit for bit, on two calls in a row (the second smaller, so stale
partial sums of the first would show), at the full ring-share size, and
refuses a size beyond the scratch before queueing anything."""
import numpy as np
import pytest

import nbody_amd
from conftest import DeviceArray

pytestmark = pytest.mark.gpu


def records(n, seed):
    m, pos, _, _ = nbody_amd.init_system("plummer", n, seed, [1.0, 20.0])
    return np.stack([pos[0], pos[1], pos[2], m], axis=1)


def both(s, form, eps, pi, pj, j_block):
    ni, nj = pi.shape[0], pj.shape[0]
    dpi, dpj = DeviceArray(pi), DeviceArray(pj)
    out_a, out_b = DeviceArray(np.full((ni, 4), 7.0)), DeviceArray(np.full((ni, 4), -7.0))
    try:
        s.accel_block_device(form, eps, ni, dpi.addr, nj, dpj.addr, j_block, out_a.addr)
        s.accel_block_device_noalloc(form, eps, ni, dpi.addr, nj, dpj.addr, j_block, out_b.addr)
        s.synchronize()
        return out_a.get(), out_b.get()
    finally:
        for b in (dpi, dpj, out_a, out_b):
            b.free()


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_noalloc_matches_allocating_entry_point_twice(form):
    with nbody_amd.Solver(0) as s:
        pj = records(20000, 3)
        a, b = both(s, form, 0.01, pj[:6000], pj, 4096)            # 5 j-blocks, many splits each
        assert np.array_equal(a, b)
        a2, b2 = both(s, form, 0.01, pj[:1000], pj[:5000], 0)      # smaller: one block
        assert np.array_equal(a2, b2)
        assert not np.array_equal(a2[:, :3], a[:1000, :3])          # a different problem, really recomputed


def test_noalloc_full_ring_share():
    """The 131072 x 1048576 ring share in j-blocks of 131072: exactly the
    static scratch's 256 MiB of partial sums."""
    with nbody_amd.Solver(0) as s:
        pj = records(1048576, 5)
        a, b = both(s, "fp64", 0.005, pj[:131072], pj, 131072)
        assert np.array_equal(a, b)


def test_noalloc_refuses_beyond_scratch_and_empty_j():
    with nbody_amd.Solver(0) as s:
        pj = records(2048, 4)
        dpj, out = DeviceArray(pj), DeviceArray(np.ones((2048, 4)))
        try:
            # 2048 i-particles against 2^21 j in blocks of 64: 32768 blocks x splits > 256 MiB
            big = DeviceArray(np.zeros((1 << 21, 4)))
            with pytest.raises(nbody_amd.NBodyError, match="static scratch"):
                s.accel_block_device_noalloc("fp64", 0.0, 2048, out.addr, 1 << 21, big.addr, 64, out.addr)
            big.free()
            s.accel_block_device_noalloc("fp64", 0.0, 2048, dpj.addr, 0, dpj.addr, 0, out.addr)
            s.synchronize()
            assert np.array_equal(out.get(), np.zeros((2048, 4)))
        finally:
            for b in (dpj, out):
                b.free()
