"""Snapshot files in the reference's layout (NBody_system::save_particles,
CPU/nbody_system.cpp:133-152): byte layout against a numpy construction,
round trips, the reference's own writer read back, and error handling.
Host-only: no GPU needed."""
import os

import numpy as np
import pytest

import nbody_amd


def state(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.5, 1.5, n), rng.normal(size=(3, n)), rng.normal(size=(3, n)), rng.normal(size=(3, n))


def test_layout_is_the_reference_frame(tmp_path):
    m, pos, vel, f = state(5, 1)
    p = str(tmp_path / "s.dat")
    nbody_amd.snapshot_write(p, m, pos, vel, f, t=0.25)
    expect = np.concatenate([[0.25], m, pos[0], pos[1], pos[2], vel[0], vel[1], vel[2], f[0], f[1], f[2]])
    assert np.array_equal(np.fromfile(p, dtype=np.float64), expect)


def test_round_trip_frames_and_append(tmp_path):
    n = 37
    p = str(tmp_path / "s.dat")
    frames = [state(n, k) for k in range(3)]
    for k, (m, pos, vel, f) in enumerate(frames):
        nbody_amd.snapshot_write(p, m, pos, vel, f if k != 1 else None, t=0.1 * k, append=k > 0)
    assert nbody_amd.snapshot_frames(p, n) == 3
    for k, (m, pos, vel, f) in enumerate(frames):
        t, m2, p2, v2, f2 = nbody_amd.snapshot_read(p, n, k)
        assert t == 0.1 * k and np.array_equal(m2, m) and np.array_equal(p2, pos) and np.array_equal(v2, vel)
        assert np.array_equal(f2, f if k != 1 else np.zeros_like(f))
    # a new write without append starts the file again
    nbody_amd.snapshot_write(p, *frames[0][:3])
    assert nbody_amd.snapshot_frames(p, n) == 1


def test_reads_the_reference_writer(ref, tmp_path, monkeypatch):
    n, dt = 64, 0.01
    m, pos, vel, f = state(n, 7)
    monkeypatch.chdir(tmp_path)
    ref.save_particles(m, pos, vel, f, dt, 3)
    path = str(tmp_path / f"NBody_{n}.dat")
    assert nbody_amd.snapshot_frames(path, n) == 3
    for k in range(3):
        t, m2, p2, v2, f2 = nbody_amd.snapshot_read(path, n, k)
        assert t == k * dt
        assert np.array_equal(m2, m) and np.array_equal(p2, pos) and np.array_equal(f2, f)
        assert np.array_equal(v2, vel * (1 + k))


def test_errors(tmp_path):
    p = str(tmp_path / "s.dat")
    m, pos, vel, f = state(10, 2)
    nbody_amd.snapshot_write(p, m, pos, vel, f)
    with pytest.raises(nbody_amd.NBodyError):
        nbody_amd.snapshot_frames(p, 11)             # not a whole number of frames for N = 11
    with pytest.raises(nbody_amd.NBodyError):
        nbody_amd.snapshot_read(p, 10, 1)            # past the end
    with pytest.raises(nbody_amd.NBodyError):
        nbody_amd.snapshot_frames(str(tmp_path / "missing.dat"), 10)
    with pytest.raises(nbody_amd.NBodyError):
        nbody_amd.snapshot_write(str(tmp_path / "no" / "dir.dat"), m, pos, vel)
