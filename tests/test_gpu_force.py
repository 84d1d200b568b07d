"""GPU force parity: the HIP kernels through the C ABI against the oracle and
the reference's golden fixtures.

Bars: FP64 form within 1e-12 relative per particle of the compensated oracle
(the kernel's rsq + cubic correction is accurate to a few ulp; the sums run in
a different order); mixed form within 2e-5 relative per particle (the
reference's own FP32 form reaches ~1e-5 on the jittered fixture), bitwise
reproducible across runs.
"""
import numpy as np
import pytest

import nbody_amd
from conftest import golden, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 2e-5


def upload(solver, m, pos, vel=None):
    solver.set_state(m, pos, vel)


@pytest.mark.parametrize("n", [64, 257, 2049])
def test_ring_fixture_forces(solver, n):
    g = golden(f"ring_{n}.npz")
    upload(solver, g["m"], g["pos"])
    f = solver.compute_forces("fp64", 0.0)
    scale = np.max(np.abs(g["force_avx2"]))
    assert np.max(np.abs(f - g["force_avx2"])) < 1e-12 * scale
    assert np.max(np.abs(f - g["force_avx512"])) < 1e-12 * scale
    assert np.max(np.abs(f - g["force_analytic"])) < 1e-12 * scale


@pytest.mark.parametrize("n", [64, 300])
def test_jitter_fixture_forces(solver, n):
    g = golden(f"jitter_{n}.npz")
    upload(solver, g["m"], g["pos"])
    f = solver.compute_forces("fp64", 0.0)
    assert rel_err(f, g["force_avx2"]) < FP64_TOL
    fm = solver.compute_forces("mixed", 0.0)
    err_mixed = rel_err(fm, g["force_avx2"])
    err_ref_sp = rel_err(g["force_sp"], g["force_avx2"])
    assert err_mixed < MIXED_TOL
    assert err_mixed < 3 * err_ref_sp + 1e-6   # no worse than the reference's FP32 form


@pytest.mark.parametrize("name", ["jitter_64.npz", "jitter_300.npz", "ring_257.npz", "ring_2049.npz"])
def test_fp32_form_at_the_reference_sp_error_level(solver, name, capsys):
    """NB_FORM_FP32 (the CLI's -p sp) against the reference's own FP32 kernel
    compute_forces_avx2_blocked_sp: both measured against the reference's
    FP64 result, the GPU form's error is of the same order: no more than a
    few times the reference's, and on the jittered systems above what the
    mixed form reaches with its FP64 differences."""
    g = golden(name)
    upload(solver, g["m"], g["pos"])
    # the ring's centre (the last particle) feels a force that is zero up to rounding
    k = slice(0, -1) if name.startswith("ring") else slice(None)
    ref64 = g["force_avx2"][:, k]
    err32 = rel_err(solver.compute_forces("fp32", 0.0)[:, k], ref64)
    err_ref = rel_err(g["force_sp"][:, k], ref64)
    err_mixed = rel_err(solver.compute_forces("mixed", 0.0)[:, k], ref64)
    assert solver.kernel_info()["kernel"] == "tile"
    with capsys.disabled():
        print(f"\n[fp32 {name}] fp32 {err32:.2e}, reference sp {err_ref:.2e}, mixed {err_mixed:.2e}")
    assert err32 < 3 * err_ref
    if name.startswith("jitter"):   # on the symmetric ring every form's error is the positions' rounding
        assert err32 > err_mixed


@pytest.mark.parametrize("kind,n,eps", [("sphere", 2048, 1e-3), ("plummer", 4096, 0.01), ("plummer", 1000, 0.0),
                                        ("sphere", 777, 0.05)])
def test_forces_vs_oracle(solver, orc, kind, n, eps):
    m, pos, vel, _ = nbody_amd.init_system(kind, n, 17)
    upload(solver, m, pos, vel)
    ref = orc.accel(m, pos, eps)
    a = solver.accel("fp64", eps)
    assert rel_err(a, ref) < FP64_TOL
    for form in ("mixed", "mixed-raw"):
        assert rel_err(solver.accel(form, eps), ref) < MIXED_TOL, form
    assert rel_err(solver.accel("fp32", eps), ref) < 1e-3


@pytest.mark.parametrize("eps", [0.0, 1e-3, 0.05])
def test_softened_ring_closed_form(solver, orc, eps):
    n = 513
    m, pos, vel, _ = nbody_amd.init_system("ring", n, 0, [1.0, 2 * np.pi, 1e2])
    upload(solver, m, pos)
    exact = orc.ring_accel_exact(n, 1.0, m[0], m[-1], eps)
    a = solver.accel("fp64", eps)
    # neighbours at +-i nearly cancel (condition number ~100): FP64 sums in a
    # different order land within ~3e-12 of the long-double closed form
    assert rel_err(a[:, :-1], exact[:, :-1]) < 1e-11


@pytest.mark.parametrize("n", [1, 2, 3, 7, 63, 64, 65, 255, 256, 257, 1023, 1025])
def test_ragged_sizes(solver, orc, n):
    m, pos, _, _ = nbody_amd.init_system("plummer", n, 5)
    m = m * np.linspace(0.5, 2.0, n)
    upload(solver, m, pos)
    for eps in (0.0, 0.02):
        a = solver.accel("fp64", eps)
        ref = orc.accel(m, pos, eps)
        if n == 1:
            assert np.all(a == 0.0)
        else:
            assert rel_err(a, ref) < FP64_TOL


def test_empty_system(solver):
    upload(solver, np.zeros(0), np.zeros((3, 0)))
    assert solver.accel("fp64", 0.0).shape == (3, 0)
    solver.step(3, 1e-3, 2, "fp64", 0.0)
    assert solver.energy(0.0) == (0.0, 0.0)


def test_coincident_particles_and_zero_mass(solver, orc):
    # duplicated positions and massless tracers: zero-separation pairs add nothing
    m, pos, _, _ = nbody_amd.init_system("sphere", 300, 8)
    pos[:, 150:160] = pos[:, 10:20]
    m[200:210] = 0.0
    upload(solver, m, np.ascontiguousarray(pos))
    for eps in (0.0, 0.01):
        a = solver.accel("fp64", eps)
        assert np.all(np.isfinite(a))
        assert rel_err(a, orc.accel(m, pos, eps)) < FP64_TOL
        am = solver.accel("mixed", eps)
        assert np.all(np.isfinite(am))


def test_bitwise_reproducible(solver):
    m, pos, _, _ = nbody_amd.init_system("plummer", 50000, 3)
    upload(solver, m, pos)
    for form in ("fp64", "mixed"):
        a1 = solver.accel(form, 0.01)
        a2 = solver.accel(form, 0.01)
        assert np.array_equal(a1, a2), form


def test_unequal_masses_momentum(solver, orc):
    # full-size property: sum_i m_i a_i = 0 (Newton's third law) at N = 262144
    # with masses apart (the grouped kernel's general path and the finish
    # that takes one particle from each end), and sampled rows against the
    # oracle
    n = 262144
    _, pos, _, _ = nbody_amd.init_system("plummer", n, 99)
    m = np.random.default_rng(99).uniform(0.5, 1.5, n) / n
    upload(solver, m, pos)
    rows = np.random.default_rng(8).choice(n, 128, replace=False)
    rows[0], rows[1] = 0, n - 1
    pj = np.stack([pos[0], pos[1], pos[2], m], axis=1)
    ref = orc.accel_block(pj[rows], pj, 0.005).T
    for form, tol, ptol in (("fp64", 1e-13, FP64_TOL), ("mixed", 1e-6, MIXED_TOL)):
        a = solver.accel(form, 0.005)
        name = solver.kernel_info()["name"]
        assert name.startswith("k_accel_symg<") and "equal" not in name, name
        tot = np.abs((m * a).sum(axis=1))
        scale = (m * np.abs(a)).sum(axis=1)
        assert np.all(tot < tol * scale), (form, tot / scale)
        assert rel_err(a[:, rows], ref) < ptol, form


def test_full_size_mixed_vs_fp64(solver):
    n = 262144
    m, pos, _, _ = nbody_amd.init_system("plummer", n, 99)
    upload(solver, m, pos)
    a64 = solver.accel("fp64", 0.005)
    am = solver.accel("mixed", 0.005)
    assert rel_err(am, a64) < MIXED_TOL


def test_accel_block_vs_oracle(solver, orc):
    m, pos, _, _ = nbody_amd.init_system("plummer", 6000, 12)
    pj = np.stack([pos[0], pos[1], pos[2], m * np.linspace(0.5, 1.5, 6000)], axis=1)
    pi = pj[1000:2500]                       # i is a subset of j: self pairs present
    ref = orc.accel_block(pi, pj, 0.01)
    for jb in (0, 1000, 2048):
        acc = solver.accel_block(pi, pj, "fp64", 0.01, j_block=jb)
        assert rel_err(acc.T, ref.T) < FP64_TOL, jb
    accm = solver.accel_block(pi, pj, "mixed", 0.01, j_block=1500)
    assert rel_err(accm.T, ref.T) < MIXED_TOL
    # eps = 0 with i inside j: the self pair must be masked
    acc0 = solver.accel_block(pi, pj, "fp64", 0.0)
    assert rel_err(acc0.T, orc.accel_block(pi, pj, 0.0).T) < FP64_TOL


def test_accel_block_empty(solver):
    pj = np.zeros((0, 4))
    pi = np.ones((5, 4))
    assert np.all(solver.accel_block(pi, pj, "fp64", 0.0) == 0.0)
    assert solver.accel_block(np.zeros((0, 4)), pi, "fp64", 0.0).shape == (0, 3)


def test_argument_errors_on_device(solver):
    with pytest.raises(nbody_amd.NBodyError):
        solver.accel("fp64", -1.0)
    with pytest.raises(nbody_amd.NBodyError):
        solver.accel(9, 0.0)
    with pytest.raises(nbody_amd.NBodyError):
        solver.step(1, 1e-3, coeffs=(np.array([0.5]), np.array([]), 1))


# --- both force kernels on the resident system ---------------------------------

@pytest.mark.parametrize("kernel", ["tile", "symmetric"])
@pytest.mark.parametrize("n", [2, 3, 100, 257, 513, 4097, 20000])
@pytest.mark.parametrize("eps", [0.0, 0.01])
def test_kernels_vs_oracle(solver, orc, kernel, n, eps):
    m, pos, _, _ = nbody_amd.init_system("plummer", n, 31)
    m = m * np.linspace(0.3, 3.0, n)   # unequal masses: a swapped m_i / m_j shows up
    solver.set_kernel(kernel)
    try:
        solver.set_state(m, pos)
        ref = orc.accel(m, pos, eps)
        assert rel_err(solver.accel("fp64", eps), ref) < FP64_TOL
        assert rel_err(solver.accel("mixed", eps), ref) < MIXED_TOL
        assert rel_err(solver.accel("mixed-raw", eps), ref) < MIXED_TOL
    finally:
        solver.set_kernel("auto")


@pytest.mark.parametrize("kernel", ["tile", "symmetric"])
def test_mixed_has_no_scale_bias(solver, orc, kernel):
    """The Newton step's factor rides on the FP32 masses (c = 1/(fl32(5/3) - 1),
    not 3/2): the mean projected error (a_mixed - a_ref).a_ref / |a_ref|^2 over
    20000 particles stays at the level the FP32 rounding of m c leaves (m = 2^-15
    makes that rounding the one of c alone, +2e-8), while the factor 3/2 would
    shrink every force by 6e-8."""
    n = 20000
    _, pos, _, _ = nbody_amd.init_system("plummer", n, 12)
    m = np.full(n, 2.0 ** -15)
    ref = orc.accel(m, pos, 0.01)
    solver.set_kernel(kernel)
    try:
        solver.set_state(m, pos)
        am = solver.accel("mixed", 0.01)
    finally:
        solver.set_kernel("auto")
    bias = float(np.mean(np.sum((am - ref) * ref, axis=0) / np.sum(ref * ref, axis=0)))
    assert abs(bias) < 4e-8, bias


@pytest.mark.parametrize("n", [40000, 131072 + 77])
def test_symmetric_matches_tile_large(solver, n):
    # K = 4 and K = 8 blocks with a ragged last block; both kernels agree to the FP64 bar
    m, pos, _, _ = nbody_amd.init_system("plummer", n, 8)
    solver.set_state(m, pos)
    solver.set_kernel("tile")
    a_tile = solver.accel("fp64", 0.005)
    solver.set_kernel("symmetric")
    a_sym = solver.accel("fp64", 0.005)
    a_sym2 = solver.accel("fp64", 0.005)
    solver.set_kernel("auto")
    assert rel_err(a_sym, a_tile) < FP64_TOL
    assert np.array_equal(a_sym, a_sym2)


def test_symmetric_coincident_eps0(solver, orc):
    # zero-separation pairs inside and across blocks must not produce NaN at eps = 0
    m, pos, _, _ = nbody_amd.init_system("sphere", 1500, 8)
    pos[:, 700:720] = pos[:, 10:30]      # same block for K = 4 (256) and a far block
    pos[:, 1400:1402] = pos[:, 1398:1400]
    pos = np.ascontiguousarray(pos)
    solver.set_kernel("symmetric")
    try:
        solver.set_state(m, pos)
        for form in ("fp64", "mixed"):
            a = solver.accel(form, 0.0)
            assert np.all(np.isfinite(a)), form
        assert rel_err(solver.accel("fp64", 0.0), orc.accel(m, pos, 0.0)) < FP64_TOL
    finally:
        solver.set_kernel("auto")


@pytest.mark.parametrize("equal", [True, False])
def test_grouped_coincident_eps0(solver, orc, equal):
    """The grouped kernel (N >= 131072) at eps = 0 with zero-separation pairs:
    inside one block, across the waves of one group (blocks 1 and 2 for the
    FP64 form's 512-particle blocks), across groups and in the padded last
    block; the equal-mass variant and the general path (one mass changed by
    2^-50).  Such pairs contribute nothing, as in the oracle."""
    n = 131072 + 131
    m, pos, _, _ = nbody_amd.init_system("plummer", n, 23)
    pairs = [(50, 100), (600, 1500), (10, 70000), (n - 2, n - 1), (5, n - 3)]
    for a, b in pairs:
        pos[:, b] = pos[:, a]
    pos = np.ascontiguousarray(pos)
    if not equal:
        m = m.copy()
        m[n // 3] *= 1.0 + 2.0 ** -50
    solver.set_state(m, pos)
    rows = np.random.default_rng(3).choice(n, 200, replace=False)
    rows = np.unique(np.concatenate([rows, np.array(pairs).ravel()]))
    pj = np.stack([pos[0], pos[1], pos[2], m], axis=1)
    ref = orc.accel_block(pj[rows], pj, 0.0).T
    for form, tol in (("fp64", FP64_TOL), ("mixed", MIXED_TOL)):
        a = solver.accel(form, 0.0)
        name = solver.kernel_info()["name"]
        assert name.startswith("k_accel_symg<") and name.endswith("equal masses>") == equal, name
        assert np.all(np.isfinite(a)), form
        assert rel_err(a[:, rows], ref) < tol, form


def test_symmetric_beyond_16_gib_of_partials(solver):
    """N = 1048576 needs 35 GB (FP64) / 24 GB (mixed) of partial sums: the
    symmetric kernel still runs (the buffer is sized against the free HBM)
    and agrees with the tile kernel; its time shows it did half the pair
    work."""
    n = 1048576
    m, pos, _, _ = nbody_amd.init_system("plummer", n, 31)
    solver.set_state(m, pos)
    for form, tol in (("fp64", FP64_TOL), ("mixed", 2e-5)):
        timed = {}
        out = {}
        for kernel in ("tile", "auto"):
            solver.set_kernel(kernel)
            solver.timing(True)
            solver.timing_reset()
            out[kernel] = solver.accel(form, 0.005)
            timed[kernel] = solver.timing_read()[0]
            solver.timing(False)
        solver.set_kernel("auto")
        assert rel_err(out["auto"], out["tile"]) < tol, form
        assert timed["auto"] < 0.8 * timed["tile"], (form, timed)


def test_accel_block_four_rows_per_lane(solver, orc):
    # 131372 i-rows: on 256 CUs the tile kernel holds 4 i-particles per lane
    # (from 512 rows per CU) and the last workgroup is partly empty.
    m, pos, _, _ = nbody_amd.init_system("plummer", 131372, 21)
    pi = np.stack([pos[0], pos[1], pos[2], m], axis=1)
    pj = pi[::64].copy()
    ref = orc.accel_block(pi, pj, 0.005)
    acc = solver.accel_block(pi, pj, "fp64", 0.005, j_block=1024)
    assert rel_err(acc.T, ref.T) < FP64_TOL
    accm = solver.accel_block(pi, pj, "mixed", 0.005)
    assert rel_err(accm.T, ref.T) < MIXED_TOL


def test_full_size_vs_oracle_sampled(solver, orc):
    """The bench workload (Plummer N = 262144, eps = 0.005, symmetric kernel):
    FP64 and mixed accelerations against the oracle on 1024 sampled rows (the
    oracle sums each sampled row over all N particles)."""
    n = 262144
    m, pos, _, _ = nbody_amd.init_system("plummer", n, 5, [1.0, 20.0])
    upload(solver, m, pos)
    rows = np.random.default_rng(4).choice(n, 1024, replace=False)
    pj = np.stack([pos[0], pos[1], pos[2], m], axis=1)
    ref = orc.accel_block(pj[rows], pj, 0.005).T
    assert rel_err(solver.accel("fp64", 0.005)[:, rows], ref) < FP64_TOL
    assert rel_err(solver.accel("mixed", 0.005)[:, rows], ref) < MIXED_TOL


def test_large_n_grouped_kernel_sampled(solver, orc):
    """N = 2^21 + 37 (a ragged last block): the grouped symmetric kernel with
    ~26 GB (FP64) / ~17 GB (mixed) of partial sums and slot indices past
    2^31 elements, FP64 and mixed, against the oracle on 256 sampled rows."""
    n = (1 << 21) + 37
    m, pos, _, _ = nbody_amd.init_system("plummer", n, 17, [1.0, 20.0])
    upload(solver, m, pos)
    rows = np.random.default_rng(6).choice(n, 256, replace=False)
    rows[0], rows[1] = 0, n - 1   # the first and the last (padded) block
    pj = np.stack([pos[0], pos[1], pos[2], m], axis=1)
    ref = orc.accel_block(pj[rows], pj, 0.005).T
    a64 = solver.accel("fp64", 0.005)
    assert solver.kernel_info()["kernel"] == "symmetric"
    assert rel_err(a64[:, rows], ref) < FP64_TOL
    assert rel_err(solver.accel("mixed", 0.005)[:, rows], ref) < MIXED_TOL


def test_ring_share_full_size_sampled(solver, orc):
    """One rank's share of the ring at full size: the first 131072 of a
    Plummer N = 1048576 against all of them in j-blocks of 131072, FP64 and
    mixed, against the oracle on 512 sampled rows."""
    ni, nj = 131072, 1048576
    m, pos, _, _ = nbody_amd.init_system("plummer", nj, 5, [1.0, 20.0])
    pj = np.stack([pos[0], pos[1], pos[2], m], axis=1)
    rows = np.random.default_rng(7).choice(ni, 512, replace=False)
    ref = orc.accel_block(pj[rows], pj, 0.005).T
    acc = solver.accel_block(pj[:ni], pj, "fp64", 0.005, j_block=131072)
    assert rel_err(acc.T[:, rows], ref) < FP64_TOL
    accm = solver.accel_block(pj[:ni], pj, "mixed", 0.005, j_block=131072)
    assert rel_err(accm.T[:, rows], ref) < MIXED_TOL


@pytest.mark.parametrize("form,n,kernel,name", [
    ("fp64", 2048, "tile", "k_accel<FP64, IPT=1>"),
    ("mixed", 16383, "tile", "k_accel<MIXED_NR, IPT=1>"),
    ("fp64", 16384, "symmetric", "k_accel_sym<FP64, K=4, equal masses>"),
    ("mixed-raw", 40000, "symmetric", "k_accel_sym<MIXED, K=4, equal masses>"),
    ("fp64", 262144, "symmetric", "k_accel_symg<FP64, K=8, W=4, equal masses>"),
    ("mixed", 262144, "symmetric", "k_accel_symg<MIXED_NR, K=6, W=4, equal masses>"),
])
def test_kernel_info_names_the_kernel_that_ran(solver, form, n, kernel, name):
    """nb_kernel_info_read reports the kernel the library chose for the last
    force evaluation (the bench labels its rooflines with it)."""
    m, pos, _, _ = nbody_amd.init_system("plummer", n, 3, [1.0, 20.0])
    upload(solver, m, pos)
    assert solver.kernel_info()["kernel"] == "auto"   # nothing evaluated since the state was set
    solver.accel(form, 0.01)
    k = solver.kernel_info()
    assert (k["kernel"], k["name"], k["form"]) == (kernel, name, form)
    if kernel == "symmetric":
        assert k["block"] == 64 * k["per_lane"] and k["scratch_bytes"] > 0
    solver.set_kernel("tile")
    try:
        solver.accel(form, 0.01)
        assert solver.kernel_info()["kernel"] == "tile"
    finally:
        solver.set_kernel("auto")


@pytest.mark.parametrize("n", [40000 + 13, 131072 + 77])
@pytest.mark.parametrize("form", ["fp64", "mixed", "mixed-raw"])
def test_equal_mass_sums_match_the_general_path(solver, orc, n, form):
    """Every mass equal: the symmetric kernels leave the mass out of the pair
    terms and the finish multiplies it in.  Against the same system with one
    mass changed by 2^-50 (the general path: masses in every pair term) and
    the oracle on sampled rows; a ragged last block (padding) at both sizes
    (item kernel K = 4 at 40013, grouped kernel at 131149)."""
    m, pos, _, _ = nbody_amd.init_system("plummer", n, 12, [1.0, 20.0])
    tol = FP64_TOL if form == "fp64" else MIXED_TOL
    solver.set_state(m, pos)
    a_eq = solver.accel(form, 0.005)
    assert solver.kernel_info()["name"].endswith("equal masses>")
    m2 = m.copy()
    m2[n // 2] *= 1.0 + 2.0 ** -50
    assert m2[n // 2] != m[0]
    solver.set_state(m2, pos)
    a_gen = solver.accel(form, 0.005)
    assert "equal" not in solver.kernel_info()["name"]
    assert rel_err(a_eq, a_gen) < tol   # mixed: two FP32 roundings of each term, each within its bar
    rows = np.random.default_rng(n).choice(n, 256, replace=False)
    rows[0] = n - 1   # a particle of the padded last block
    pj = np.stack([pos[0], pos[1], pos[2], m], axis=1)
    ref = orc.accel_block(pj[rows], pj, 0.005).T
    assert rel_err(a_eq[:, rows], ref) < tol


@pytest.mark.parametrize("form", ["fp64", "mixed", "fp32"])
def test_tile_kernel_equal_masses(solver, orc, form):
    """The tile kernel with every j mass equal leaves the mass out of the
    pair terms (full j-tiles) and puts it into the final scale: against the
    same j set with one mass changed by 2^-50 (the general path) and the
    oracle; j-blocks with a partial last tile, the resident system too."""
    tol = {"fp64": FP64_TOL, "mixed": MIXED_TOL, "fp32": 1e-3}[form]
    m, pos, _, _ = nbody_amd.init_system("plummer", 5000, 14, [1.0, 20.0])
    pj = np.stack([pos[0], pos[1], pos[2], m], axis=1)
    pi = pj[::7].copy()
    ref = orc.accel_block(pi, pj, 0.005).T
    a_eq = solver.accel_block(pi, pj, form, 0.005, j_block=1300).T
    pj2 = pj.copy()
    pj2[2500, 3] *= 1.0 + 2.0 ** -50
    a_gen = solver.accel_block(pi, pj2, form, 0.005, j_block=1300).T
    assert rel_err(a_eq, ref) < tol and rel_err(a_gen, ref) < tol
    assert rel_err(a_eq, a_gen) < tol
    solver.set_kernel("tile")
    try:
        solver.set_state(m, pos)
        a_res = solver.accel(form, 0.005)
    finally:
        solver.set_kernel("auto")
    assert rel_err(a_res[:, ::7], ref) < tol
