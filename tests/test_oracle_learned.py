"""The learned-extrapolator oracle pinned on the CPU: backward pass against
central finite differences, Adam against its closed form, the cubic
extrapolation against polynomials, the features' scale invariance and the
bfloat16 rounding against known bit patterns."""
import numpy as np
import pytest


def random_params(orc, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.empty(orc.mlp_param_count())
    sizes = [(128, 15), (128,), (128, 128), (128,), (128, 128), (128,), (3, 128), (3,)]
    off = 0
    for s in sizes:
        k = int(np.prod(s))
        fan_in = s[1] if len(s) == 2 else 8
        p[off:off + k] = rng.uniform(-1, 1, k) * scale / np.sqrt(fan_in)
        off += k
    assert off == p.shape[0] == 35459
    return p


def test_bf16_rounding(orc):
    assert orc.bf16(1.0) == 1.0
    assert orc.bf16(1.0 + 2 ** -8) == 1.0            # tie -> even
    assert orc.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6   # tie -> even (up)
    assert orc.bf16(1.0 + 2 ** -7) == 1.0 + 2 ** -7     # representable
    assert orc.bf16(-3.14159) == -3.140625


def test_backward_matches_finite_differences(orc):
    p = random_params(orc, 3)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, (6, 15))
    t = rng.uniform(-1, 1, (6, 3))
    g, loss = orc.mlp_grad(p, x, t, bf16=False)
    y = orc.mlp_forward(p, x, bf16=False)
    assert abs(loss - 0.5 * np.sum((y - t) ** 2) / 6) < 1e-14
    idx = np.concatenate([rng.choice(1920, 6), 1920 + rng.choice(128, 3), 2048 + rng.choice(16384, 6),
                          35072 + rng.choice(384, 4), [35456, 35457, 35458]])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (orc.mlp_grad(pp, x, t, bf16=False)[1] - orc.mlp_grad(pm, x, t, bf16=False)[1]) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 * max(1.0, abs(g[k])) + 1e-9, (k, fd, g[k])


def test_adam_closed_form(orc):
    n = 5
    p = np.linspace(-1, 1, n)
    g = np.array([0.5, -0.25, 1e-3, 0.0, 2.0])
    m, v = np.zeros(n), np.zeros(n)
    p0 = p.copy()
    orc.adam(p, g, m, v, 1e-2, 1)
    # first step: m_hat = g, v_hat = g^2 -> p -= lr * sign(g) (up to eps)
    expect = p0 - 1e-2 * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(p, expect, rtol=0, atol=1e-12)


@pytest.mark.parametrize("b1,b2", [(0.9, 0.999), (0.5, 0.75)])
def test_adam_two_steps_closed_form(orc, b1, b2):
    """Two steps with different gradients fix beta1, beta2 and the
    t-dependence of the bias correction: after g1 then g2 the corrected
    moments are (b1 g1 + g2) / (1 + b1) and (b2 g1^2 + g2^2) / (1 + b2)."""
    lr, eps = 1e-2, 1e-8
    p0 = np.linspace(-1, 1, 6)
    g1 = np.array([0.5, -0.25, 1e-3, 0.0, 2.0, -1.0])
    g2 = np.array([-0.5, 0.75, 3e-3, 1.0, 2.0, 0.25])
    p, m, v = p0.copy(), np.zeros(6), np.zeros(6)
    orc.adam(p, g1, m, v, lr, 1, b1, b2, eps)
    orc.adam(p, g2, m, v, lr, 2, b1, b2, eps)
    p1 = p0 - lr * g1 / (np.abs(g1) + eps)
    mh = (b1 * g1 + g2) / (1 + b1)
    vh = (b2 * g1 ** 2 + g2 ** 2) / (1 + b2)
    np.testing.assert_allclose(p, p1 - lr * mh / (np.sqrt(vh) + eps), rtol=0, atol=1e-14)
    np.testing.assert_allclose(m, (1 - b1) * (b1 * g1 + g2), rtol=0, atol=1e-15)
    np.testing.assert_allclose(v, (1 - b2) * (b2 * g1 ** 2 + g2 ** 2), rtol=0, atol=1e-15)
    # without the bias correction of step 2 the update would differ
    wrong = p1 - lr * ((1 - b1) * (b1 * g1 + g2) / (1 - b1)) / (np.sqrt((1 - b2) * (b2 * g1 ** 2 + g2 ** 2) / (1 - b2)) + eps)
    assert np.max(np.abs(p - wrong)) > 1e-4


def single_path_params(orc):
    """Weights with one active path x[2] -> h1[0] -> h2[5] -> h3[7] -> y[1]
    (W_k row-major [out][in]), biases on that path and on y."""
    P = np.zeros(orc.mlp_param_count())
    W1, b1, W2, b2, W3, b3, W4, b4 = 0, 1920, 2048, 18432, 18560, 34944, 35072, 35456
    P[W1 + 0 * 15 + 2] = 1.5
    P[b1 + 0] = -0.25
    P[W2 + 5 * 128 + 0] = -2.0
    P[b2 + 5] = 0.5
    P[W3 + 7 * 128 + 5] = 3.0
    P[b3 + 7] = -1.0
    P[W4 + 1 * 128 + 7] = 0.75
    P[b4:b4 + 3] = [0.125, -0.0625, 0.25]
    return P


def single_path_expect(x):
    silu = lambda z: z / (1.0 + np.exp(-z))
    h1 = silu(1.5 * x[:, 2] - 0.25)
    h2 = silu(-2.0 * h1 + 0.5)
    h3 = silu(3.0 * h2 - 1.0)
    y = np.zeros((x.shape[0], 3))
    y[:, 0] = 0.125
    y[:, 1] = 0.75 * h3 - 0.0625
    y[:, 2] = 0.25
    return y


def test_mlp_forward_single_path_closed_form(orc):
    """Layout [out][in], bias placement and SiLU against a closed form: only
    one unit per layer is active; the other inputs must not leak in."""
    P = single_path_params(orc)
    rng = np.random.default_rng(9)
    x = rng.uniform(-2, 2, (7, 15))
    np.testing.assert_allclose(orc.mlp_forward(P, x, bf16=False), single_path_expect(x), rtol=0, atol=1e-14)
    # a transposed first layer ([in][out]) would read input 0 for unit 2 instead
    Pt = P.copy()
    Pt[2], Pt[2 * 128 + 0] = 0.0, 1.5
    assert np.max(np.abs(orc.mlp_forward(Pt, x, bf16=False) - single_path_expect(x))) > 1e-3


def test_cubic_extrapolation_exact_on_cubics(orc):
    rng = np.random.default_rng(1)
    coef = rng.standard_normal((4, 30))
    f = lambda t: coef[0] + coef[1] * t + coef[2] * t ** 2 + coef[3] * t ** 3
    a0, a1, a2, a3 = f(3.0), f(2.0), f(1.0), f(0.0)
    np.testing.assert_allclose(orc.poly_extrap(a0, a1, a2, a3), f(4.0), rtol=1e-12, atol=1e-12)
    # a quartic term is not reproduced: the residual is the 4th difference 24 c4
    g = lambda t: t ** 4
    assert abs(orc.poly_extrap(*[np.array([g(t)]) for t in (3.0, 2.0, 1.0, 0.0)])[0] - g(4.0)) == 24.0


def random_rotation(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def test_features_invariance_and_frame(orc):
    """Features are invariant under rotations and under scaling of a and v;
    the frame is orthonormal, right-handed, e1 along D3, e2 in the D3-D2
    plane, and turns with the data."""
    rng = np.random.default_rng(2)
    a = [rng.standard_normal((50, 3)) for _ in range(4)]
    v = rng.standard_normal((50, 3))
    x, s, cub, fr = orc.features(*a, v)
    R = random_rotation(rng)
    x2, s2, cub2, fr2 = orc.features(*[7.5 * ai @ R.T for ai in a], 0.1 * v @ R.T)
    np.testing.assert_allclose(x2, x, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(s2, 7.5 * s, rtol=1e-12)
    np.testing.assert_allclose(fr2, fr @ R.T, rtol=0, atol=1e-10)
    np.testing.assert_allclose(cub, 4 * a[0] - 6 * a[1] + 4 * a[2] - a[3], rtol=1e-12, atol=1e-12)
    d2 = a[0] - 2 * a[1] + a[2]
    d3 = a[0] - 3 * a[1] + 3 * a[2] - a[3]
    np.testing.assert_allclose(s, np.linalg.norm(d3, axis=1), rtol=1e-12)
    for i in range(50):
        E = fr[i]
        np.testing.assert_allclose(E @ E.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(E) > 0.999
        np.testing.assert_allclose(E[0], d3[i] / np.linalg.norm(d3[i]), atol=1e-12)
        assert abs(np.dot(E[2], d2[i])) < 1e-10 * np.linalg.norm(d2[i])
    assert np.all(np.abs(x[:, :13]) <= 1.0 + 1e-12) and np.all(np.abs(x[:, 13]) <= 2.0)


def test_features_reflection_flips_only_the_pseudoscalar(orc):
    rng = np.random.default_rng(4)
    a = [rng.standard_normal((20, 3)) for _ in range(4)]
    v = rng.standard_normal((20, 3))
    x, _, _, _ = orc.features(*a, v)
    M = np.diag([1.0, 1.0, -1.0])
    xm, _, _, _ = orc.features(*[ai @ M for ai in a], v @ M)
    np.testing.assert_allclose(xm[:, :14], x[:, :14], atol=1e-12)
    np.testing.assert_allclose(xm[:, 14], -x[:, 14], atol=1e-12)


def test_features_degenerate_frames(orc):
    """D2 parallel to D3 falls back to D1; all three parallel to an axis
    choice; a zero history gives zero features, scale and frame."""
    z = np.zeros((3, 3))
    x, s, cub, fr = orc.features(z, z, z, z, z)
    assert np.all(s == 0) and np.all(cub == 0) and np.all(fr == 0)
    assert np.all(x[:, :13] == 0) and np.all(x[:, 13] == -2.0) and np.all(x[:, 14] == 0)
    # samples on a line: every difference along one direction
    d = np.array([[0.0, 0.6, 0.8]])
    a = [d * c for c in (1.0, 2.0, 4.0, 8.0)]
    _, _, _, fr = orc.features(*a, d)
    E = fr[0]
    np.testing.assert_allclose(E @ E.T, np.eye(3), atol=1e-12)
    np.testing.assert_allclose(np.abs(E[0]), np.abs(d[0]), atol=1e-12)


def test_near_taper_properties(orc):
    """S = 1 inside rc/2, 0 from rc on, C2 at both joins, monotone between,
    1/2 where r^2 / rc^2 is midway (5/8)."""
    rc = 0.3
    assert orc.near_taper(0.0, rc) == 1.0 and orc.near_taper(0.15, rc) == 1.0
    assert orc.near_taper(0.3, rc) == 0.0 and orc.near_taper(0.45, rc) == 0.0
    assert abs(orc.near_taper(rc * np.sqrt(5 / 8), rc) - 0.5) < 1e-14
    # C2 at the joins: S(r0 +- h) - S(r0) = O(h^3) on the inner side, 0 outside
    for r0, side in ((0.15, 1.0), (0.3, -1.0)):
        d = [abs(orc.near_taper(r0 + side * h, rc) - orc.near_taper(r0, rc)) for h in (1e-3, 1e-4)]
        assert 0.9e3 < d[0] / d[1] < 1.1e3, (r0, d)
        assert orc.near_taper(r0 - side * 1e-3, rc) == orc.near_taper(r0, rc)
    r = np.linspace(0.15, 0.3, 301)
    s = np.array([orc.near_taper(x, rc) for x in r])
    assert np.all(np.diff(s) <= 0)


def test_near_field_two_bodies(orc):
    """Two bodies: the near field is the softened pair force times S(r), zero beyond rc."""
    rc, eps = 0.4, 0.05
    m = np.array([0.7, 0.3])
    for r in (0.1, 0.25, 0.39, 0.41):
        pos = np.array([[0.0, r], [0.0, 0.0], [0.0, 0.0]])
        a = orc.near_accel(m, pos, rc, eps)
        full = orc.accel(m, pos, eps)
        S = orc.near_taper(r, rc)
        assert np.allclose(a, S * full, rtol=1e-14, atol=0.0), r
        if r > rc:
            assert np.all(a == 0.0)


def test_near_field_large_radius_is_full_sum(orc):
    """With every pair inside rc/2 the near field is the whole softened sum."""
    rng = np.random.default_rng(8)
    n = 300
    m = rng.uniform(0.5, 1.5, n) / n
    pos = rng.uniform(-1, 1, (3, n))
    a = orc.near_accel(m, pos, 10.0, 0.01)
    full = orc.accel(m, pos, 0.01)
    assert np.max(np.abs(a - full)) < 1e-13 * np.max(np.abs(full))


def test_lagrange_weights(orc):
    """Equally spaced nodes give the backward-difference cubic; any distinct
    nodes reproduce cubics exactly and the weights sum to one."""
    assert np.allclose(orc.lagrange4([0, -1, -2, -3], 1.0), [4, -6, 4, -1], rtol=0, atol=1e-14)
    assert np.allclose(orc.lagrange4([0, -4, -8, -12], 4.0), [4, -6, 4, -1], rtol=0, atol=1e-13)
    rng = np.random.default_rng(3)
    for _ in range(20):
        t = np.sort(rng.choice(40, 4, replace=False))[::-1].astype(float)
        x = t[0] + rng.integers(1, 6)
        c = rng.normal(size=4)
        f = lambda u: c[0] + c[1] * u + c[2] * u ** 2 + c[3] * u ** 3
        w = orc.lagrange4(t, x)
        assert abs(np.sum(w) - 1) < 1e-10
        assert abs(np.dot(w, [f(u) for u in t]) - f(x)) < 1e-8 * (1 + abs(f(x)))
    # a quartic is missed by exactly its leading coefficient times the nodal polynomial
    t = np.array([0.0, -4.0, -8.0, -12.0])
    w = orc.lagrange4(t, 3.0)
    assert abs(np.dot(w, t ** 4) - 3.0 ** 4 + np.prod(3.0 - t)) < 1e-9
