"""Known-answer and invariant checks that pin the oracle where the reference
cannot (it has no softening, no Plummer or sphere ICs): closed forms,
long-double brute force, conservation laws, convergence order and
distribution statistics.  Each check is chosen so that a plausible mistake
(a dropped term, a wrong sign or index, a misplaced softening) fails it; the
`planted` tests below show that on a numpy restatement with the mistake put in.
"""
import numpy as np
import pytest
from scipy import integrate, stats

from conftest import rel_err

TWO_PI = 2 * np.pi


# ---------------------------------------------------------------------------
# forces
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("eps", [0.0, 1e-3, 0.05, 0.3])
@pytest.mark.parametrize("n", [64, 257])
def test_softened_ring_closed_form(orc, n, eps):
    m, pos, vel, f = orc.ic_ring(n, 1.0, TWO_PI, 1e2)
    exact = orc.ring_accel_exact(n, 1.0, m[0], m[-1], eps)
    a = orc.accel(m, pos, eps)
    assert rel_err(a[:, :-1], exact[:, :-1]) < 1e-12
    assert np.max(np.abs(a[:, -1])) < 1e-12 * np.max(np.abs(a))   # the centre feels nothing
    if eps == 0.0:   # and the reference's exact centripetal force (CPU/nbody_system.cpp:262-264)
        assert rel_err(a * m, f) < 1e-12


@pytest.mark.parametrize("eps", [0.0, 0.01])
@pytest.mark.parametrize("kind", ["sphere", "plummer"])
def test_long_double_brute_force(orc, kind, eps):
    m, pos, _ = orc.ic_sphere(64, 5) if kind == "sphere" else orc.ic_plummer(64, 5)
    m = m * np.linspace(0.2, 3.0, 64)   # unequal masses expose an m_i / m_j swap
    a = orc.accel(m, pos, eps)
    ald = orc.accel(m, pos, eps, long_double=True)
    assert rel_err(a, ald) < 1e-14


def test_momentum_conservation(orc):
    m, pos, _ = orc.ic_plummer(2000, 11)
    m = m * np.linspace(0.5, 1.5, 2000)
    a = orc.accel(m, pos, 0.01)
    total = np.abs((m * a).sum(axis=1))
    assert np.all(total < 1e-14 * np.abs(m * a).sum(axis=1))


def test_block_form_equals_full_form(orc):
    m, pos, _ = orc.ic_plummer(500, 2)
    pj = np.stack([pos[0], pos[1], pos[2], m], axis=1)
    acc = orc.accel_block(pj[100:200], pj, 0.02)
    a = orc.accel(m, pos, 0.02)
    assert np.max(np.abs(acc.T - a[:, 100:200])) <= 1e-15 * np.max(np.abs(a))


# ---------------------------------------------------------------------------
# planted mistakes: the checks above reject each of them
# ---------------------------------------------------------------------------

def numpy_accel(m, pos, eps, mistake=None):
    d = pos[:, None, :] - pos[:, :, None]            # d[:, i, j] = x_j - x_i
    r2 = (d * d).sum(axis=0)
    np.fill_diagonal(r2, np.inf)
    if mistake == "eps_outside":
        inv3 = 1.0 / (r2 ** 1.5 + eps ** 3)
    elif mistake == "eps_squared_twice":
        inv3 = (r2 + eps ** 4) ** -1.5
    else:
        inv3 = (r2 + eps * eps) ** -1.5
    np.fill_diagonal(inv3, 0.0)
    w = m[None, :] if mistake != "mass_index" else m[:, None]
    if mistake == "sign":
        w = -w
    a = (d * (w * inv3)[None]).sum(axis=2)
    if mistake == "drop_last":
        a = (d[:, :, :-1] * (w * inv3)[None, :, :-1]).sum(axis=2)
    return a


@pytest.mark.parametrize("mistake", ["sign", "eps_outside", "eps_squared_twice", "mass_index", "drop_last"])
def test_planted_mistakes_are_caught(orc, mistake):
    n, eps = 65, 0.05
    m, pos, _, _ = orc.ic_ring(n, 1.0, TWO_PI, 1e2)
    exact = orc.ring_accel_exact(n, 1.0, m[0], m[-1], eps)
    good = numpy_accel(m, pos, eps)
    assert rel_err(good[:, :-1], exact[:, :-1]) < 1e-12      # the restatement itself passes
    bad = numpy_accel(m, pos, eps, mistake)
    caught_ring = rel_err(bad[:, :-1], exact[:, :-1]) > 1e-6
    # unequal masses for the brute-force check
    m2, pos2, _ = orc.ic_plummer(40, 9)
    m2 = m2 * np.linspace(0.2, 3.0, 40)
    ald = orc.accel(m2, pos2, eps, long_double=True)
    caught_bf = rel_err(numpy_accel(m2, pos2, eps, mistake), ald) > 1e-6
    assert caught_ring or caught_bf


# ---------------------------------------------------------------------------
# integrator
# ---------------------------------------------------------------------------

def test_yoshida_coefficients(orc):
    cbrt2 = 2.0 ** (1.0 / 3.0)
    w1, w0 = 1.0 / (2.0 - cbrt2), -cbrt2 / (2.0 - cbrt2)
    c, d, lf = orc.yoshida(4)
    assert lf == 4
    np.testing.assert_allclose(c, [w1 / 2, (w0 + w1) / 2, (w0 + w1) / 2, w1 / 2], rtol=1e-15)
    np.testing.assert_allclose(d, [w1, w0, w1], rtol=1e-15)
    c, d, lf = orc.yoshida(2)
    assert lf == 2 and list(c) == [0.5, 0.5] and list(d) == [1.0]
    for order, lf_expect in ((6, 8), (8, 16)):
        c, d, lf = orc.yoshida(order)
        assert lf == lf_expect
        assert abs(c.sum() - 1) < 1e-14 and abs(d.sum() - 1) < 1e-14   # CPU/omp_solver.cpp:728-741
        assert np.array_equal(c, c[::-1]) and np.array_equal(d, d[::-1])
    # unknown orders fall back to 2 (CPU/omp_solver.cpp:703-707)
    assert orc.yoshida(3)[2] == 2


def two_body(eps, m1=0.7, m2=0.3, D=1.0):
    """Circular softened binary: omega^2 = (m1+m2)/(D^2+eps^2)^(3/2)."""
    om = np.sqrt((m1 + m2) / (D * D + eps * eps) ** 1.5)
    r1, r2 = m2 / (m1 + m2) * D, m1 / (m1 + m2) * D
    m = np.array([m1, m2])
    pos = np.array([[r1, -r2], [0.0, 0.0], [0.0, 0.0]])
    vel = np.array([[0.0, 0.0], [r1 * om, -r2 * om], [0.0, 0.0]])
    return m, np.ascontiguousarray(pos), np.ascontiguousarray(vel), TWO_PI / om


@pytest.mark.parametrize("order,steps,expect", [(2, (200, 400), 2), (4, (50, 100), 4), (6, (24, 48), 6),
                                                (8, (16, 32), 8)])
def test_convergence_order(orc, order, steps, expect):
    eps = 0.2
    m, pos, vel, P = two_body(eps)
    errs = []
    for s in steps:
        p, v = orc.leapfrog(m, pos, vel, s, P / s, order, eps)
        errs.append(np.max(np.abs(p - pos)))   # one full period returns to the start
    slope = np.log2(errs[0] / errs[1])
    assert abs(slope - expect) < 0.6, (errs, slope)


def test_time_reversibility(orc):
    m, pos, vel = orc.ic_plummer(64, 21)
    p1, v1 = orc.leapfrog(m, pos, vel, 40, 1e-3, 4, 0.05)
    p2, v2 = orc.leapfrog(m, p1, np.ascontiguousarray(-v1), 40, 1e-3, 4, 0.05)
    assert np.max(np.abs(p2 - pos)) < 1e-12
    assert np.max(np.abs(-v2 - vel)) < 1e-11


def test_ring_rotation_kat(orc):
    # a quarter revolution of the reference's ring matches the rigid rotation
    n = 129
    m, pos, vel, _ = orc.ic_ring(n, 1.0, TWO_PI, 1e5)
    s = 400
    p, v = orc.leapfrog(m, pos, vel, s, 0.25 / s, 4, 0.0)
    pp, vp = orc.ring_predict(n, 1.0, TWO_PI, 0.25)
    assert np.max(np.abs(p - pp)) < 1e-7
    assert np.max(np.abs(v - vp)) < 1e-6


# ---------------------------------------------------------------------------
# energy
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("eps", [0.0, 0.05])
def test_energy_brute_force(orc, eps):
    m, pos, vel = orc.ic_plummer(200, 4)
    m = m * np.linspace(0.3, 2.0, 200)
    T, U = orc.energy(m, pos, vel, eps)
    mp = m.astype(np.longdouble)
    pl = pos.astype(np.longdouble)
    d = pl[:, None, :] - pl[:, :, None]
    r = np.sqrt((d * d).sum(axis=0) + np.longdouble(eps) ** 2)
    iu = np.triu_indices(200, 1)
    U_ld = -(mp[:, None] * mp[None, :] / np.where(r > 0, r, 1))[iu].sum()
    T_ld = 0.5 * (mp * (vel.astype(np.longdouble) ** 2).sum(axis=0)).sum()
    assert abs(U - float(U_ld)) < 1e-15 * abs(float(U_ld))
    assert abs(T - float(T_ld)) < 1e-15 * abs(float(T_ld))


# ---------------------------------------------------------------------------
# initial conditions
# ---------------------------------------------------------------------------

def plummer_cdf(r, rmax=20.0):
    c = rmax ** 3 / (rmax ** 2 + 1) ** 1.5
    return (r ** 3 / (r ** 2 + 1) ** 1.5) / c


def test_plummer_radii_and_speeds(orc):
    m, pos, vel = orc.ic_plummer(20000, 123)
    assert abs(m.sum() - 1.0) < 1e-12
    assert np.max(np.abs((m * pos).sum(axis=1))) < 1e-14 and np.max(np.abs((m * vel).sum(axis=1))) < 1e-14
    r = np.linalg.norm(pos, axis=0)
    assert r.max() < 20.5   # truncation (plus the small centre-of-mass shift)
    assert stats.kstest(r, plummer_cdf).pvalue > 1e-3
    # speeds: q = v / v_esc(r) follows q^2 (1-q^2)^(7/2)
    q = np.linalg.norm(vel, axis=0) / (np.sqrt(2.0) * (r * r + 1) ** -0.25)
    qs = np.linspace(0, 1, 2001)
    g = qs ** 2 * (1 - qs ** 2) ** 3.5
    cdf = np.concatenate([[0], np.cumsum((g[1:] + g[:-1]) / 2 * np.diff(qs))])
    cdf /= cdf[-1]
    assert stats.kstest(np.clip(q, 0, 1), lambda x: np.interp(x, qs, cdf)).pvalue > 1e-3
    # isotropy: mean direction ~ 0
    assert np.linalg.norm((pos / r).mean(axis=1)) < 0.03


def test_plummer_energy_and_virial(orc):
    m, pos, vel = orc.ic_plummer(8192, 77)
    T, U = orc.energy(m, pos, vel, 0.0)
    # U of the truncated, renormalised profile: -int_0^rmax M(r) M'(r) / r dr
    norm = 20.0 ** 3 / (20.0 ** 2 + 1) ** 1.5
    dM = lambda r: 3 * r * r / (r * r + 1) ** 2.5 / norm
    U_th = -integrate.quad(lambda r: plummer_cdf(r) * dM(r) / r, 0, 20.0, limit=200)[0]
    assert abs(U - U_th) < 0.02 * abs(U_th)
    assert abs(U_th - (-3 * np.pi / 32)) < 0.02   # close to the untruncated -3 pi / 32
    assert 0.95 < 2 * T / abs(U) < 1.05


def test_uniform_sphere(orc):
    m, pos, vel = orc.ic_sphere(20000, 9)
    r = np.linalg.norm(pos, axis=0)
    assert r.max() < 1.0 and np.all(vel == 0.0) and abs(m.sum() - 1) < 1e-12
    assert stats.kstest(r, lambda x: np.clip(x, 0, 1) ** 3).pvalue > 1e-3
    m, pos, vel = orc.ic_sphere(4096, 9)
    T, U = orc.energy(m, pos, vel, 0.0)
    assert T == 0.0 and abs(U - (-0.6)) < 0.01   # -3 G M^2 / (5 R)


def test_ic_seed_determinism(orc):
    a = orc.ic_plummer(1000, 5)
    b = orc.ic_plummer(1000, 5)
    c = orc.ic_plummer(1000, 6)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[1], c[1])
