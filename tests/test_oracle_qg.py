"""Pins the oracle's QG model against facts fixed by the mathematics
(closed forms, polynomial exactness, identities), not against itself.

Each check is chosen so that a plausible slip (index offset, transposed
operand, wrong sign, dropped term) fails it: grids are rectangular with
dx != dy throughout.
"""
import numpy as np
import pytest

from oracle.qg import QgmOracle, build_mask, dst1, inverse_elliptic_dst

G, F0 = 9.81, 1e-4


def model(ny=24, nx=30, scheme="rk2", **kw):
    return QgmOracle(dx=10e3, dy=7e3, dt=600.0, c=2.5, f=F0, time_scheme=scheme,
                     shape=(ny, nx), **kw)


# ----------------------------------------------------------------- DST / Helmholtz
def sine_matrix(m):
    j = np.arange(1, m + 1)
    return np.sqrt(2.0 / (m + 1)) * np.sin(np.pi * np.outer(j, j) / (m + 1))


def test_dst1_is_orthonormal_sine_transform():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((13, 21))
    Sy, Sx = sine_matrix(13), sine_matrix(21)
    np.testing.assert_allclose(dst1(x), Sy @ x @ Sx, atol=1e-13)
    np.testing.assert_allclose(Sy @ Sy, np.eye(13), atol=1e-13)
    np.testing.assert_allclose(dst1(dst1(x)), x, atol=1e-13)


@pytest.mark.parametrize("k,l", [(0, 0), (3, 7), (11, 2), (21, 27)])
def test_helmholtz_eigenmodes(k, l):
    """A single DST-I mode with a zero ring is an eigenvector of h2pv's stencil
    with eigenvalue helmoltz_dst[k, l]; pv2h inverts it exactly."""
    m = model()
    ny, nx = m.ny, m.nx
    i = np.arange(ny)[:, None]
    j = np.arange(nx)[None, :]
    mode = np.sin(np.pi * i * (k + 1) / (ny - 1)) * np.sin(np.pi * j * (l + 1) / (nx - 1))
    lam = G / F0 * (2 * (np.cos(np.pi * (l + 1) / (nx - 1)) - 1) / m.dx ** 2 +
                    2 * (np.cos(np.pi * (k + 1) / (ny - 1)) - 1) / m.dy ** 2) - G * F0 / m.c ** 2
    assert m.helmoltz_dst[k, l] == pytest.approx(lam, rel=1e-14)
    zero = np.zeros((ny, nx))
    q = m.h2pv(mode, zero)
    act = m.mask == 3
    np.testing.assert_allclose(q[act], lam * mode[act], rtol=1e-11, atol=1e-11 * abs(lam))
    qfull = np.zeros((ny, nx))
    qfull[1:-1, 1:-1] = lam * mode[1:-1, 1:-1]
    np.testing.assert_allclose(m.pv2h(qfull, zero, zero), mode, atol=1e-12)


def test_inverse_elliptic_dst_is_symmetric():
    m = model(20, 26)
    rng = np.random.default_rng(2)
    a, b = rng.standard_normal((2, 18, 24))
    lhs = np.sum(inverse_elliptic_dst(a, m.helmoltz_dst) * b)
    rhs = np.sum(a * inverse_elliptic_dst(b, m.helmoltz_dst))
    assert lhs == pytest.approx(rhs, rel=1e-12)


# ----------------------------------------------------------------------- masks
def test_mask_ring_known_answer():
    expect = np.array([
        [1, 1, 1, 1, 1, 1, 1, 1, 1],
        [1, 2, 2, 2, 2, 2, 2, 2, 1],
        [1, 2, 3, 3, 3, 3, 2, 2, 1],
        [1, 2, 3, 3, 3, 3, 2, 2, 1],
        [1, 2, 3, 3, 3, 3, 2, 2, 1],
        [1, 2, 2, 2, 2, 2, 2, 2, 1],
        [1, 2, 2, 2, 2, 2, 2, 2, 1],
        [1, 1, 1, 1, 1, 1, 1, 1, 1]])
    assert np.array_equal(build_mask(8, 9), expect)


def test_mask_land_known_answer():
    ssh = np.zeros((12, 12))
    ssh[6, 6] = np.nan
    m = build_mask(12, 12, ssh)
    assert m[6, 6] == 0
    # 3x3 neighbours of land become coast (1), the 5x5 ring becomes 2
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            if di or dj:
                assert m[6 + di, 6 + dj] == 1
    assert m[4, 4] == 2 and m[8, 8] == 2 and m[6, 4] == 2
    assert m[3, 3] == 3 and m[5, 3] == 3 and m[9, 9] == 2  # row/col n-3 = 9 is a ring row


# ----------------------------------------------------- polynomial exactness
def test_h2uv_exact_for_bilinear():
    m = model(16, 18)
    y = m.dy * np.arange(m.ny)[:, None]
    x = m.dx * np.arange(m.nx)[None, :]
    a, b, c, d = 0.3, 2e-6, -5e-6, 1e-11
    h = a + b * x + c * y + d * x * y
    u, v = m.h2uv(h)
    # u is -g/f dh/dy at (y_i, x_{j-1/2}); v is g/f dh/dx at (y_{i-1/2}, x_j)
    xs = m.dx * (np.arange(m.nx) - 0.5)[None, :]
    ys = m.dy * (np.arange(m.ny) - 0.5)[:, None]
    u_ex = -G / F0 * (c + d * xs) * np.ones((m.ny, 1))
    v_ex = G / F0 * (b + d * ys) * np.ones((1, m.nx))
    np.testing.assert_allclose(u[1:-1, 1:], u_ex[1:-1, 1:], rtol=1e-10)
    np.testing.assert_allclose(v[1:, 1:-1], v_ex[1:, 1:-1], rtol=1e-10)
    assert np.all(u[0] == 0) and np.all(u[:, 0] == 0) and np.all(v[0] == 0) and np.all(v[:, 0] == 0)


def test_h2pv_exact_for_quadratic():
    m = model(16, 18)
    y = m.dy * np.arange(m.ny)[:, None]
    x = m.dx * np.arange(m.nx)[None, :]
    h = 1e-10 * x ** 2 + 3e-10 * y ** 2 + 1e-6 * x
    hb = 0.01 * np.ones_like(h)
    q = m.h2pv(h, hb)
    act = m.mask == 3
    expect = G / F0 * (2e-10 + 6e-10) - G * F0 / m.c ** 2 * h
    np.testing.assert_allclose(q[act], expect[act], rtol=1e-9)
    ring = (m.mask == 1) | (m.mask == 2)
    np.testing.assert_allclose(q[ring], -G * F0 / m.c ** 2 * 0.01, rtol=1e-14)


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_adv_exact_for_cubics(sign):
    """3rd-order upwind: exact derivative of cubic PV for either wind sign."""
    m = model(20, 22)
    y = m.dy * np.arange(m.ny)[:, None] * np.ones((1, m.nx))
    x = m.dx * np.arange(m.nx)[None, :] * np.ones((m.ny, 1))
    q = 1e-15 * x ** 3 - 2e-10 * x ** 2 + 3e-16 * y ** 3 + 1e-9 * y
    dqdx = 3e-15 * x ** 2 - 4e-10 * x
    dqdy = 9e-16 * y ** 2 + 1e-9
    U, V = 0.4 * sign, -0.25 * sign
    shp = (m.ny - 2, m.nx - 2)
    up = np.full(shp, max(U, 0)); um = np.full(shp, min(U, 0))
    vp = np.full(shp, max(V, 0)); vm = np.full(shp, min(V, 0))
    out = m.adv(up, vp, um, vm, q)
    ex = -(U * dqdx + V * dqdy)
    np.testing.assert_allclose(out[2:-2, 2:-2], ex[2:-2, 2:-2], rtol=1e-8, atol=1e-22)


# ---------------------------------------------------------------- time schemes
@pytest.mark.parametrize("scheme,poly", [
    ("Euler", lambda z: 1 + z),
    ("rk2", lambda z: 1 + z + z ** 2 / 2),
    ("rk3", lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6)])
def test_scheme_amplification(scheme, poly):
    """With rhs(q) = lam*q the step is the scheme's stability polynomial."""
    m = model(12, 12, scheme)
    lam = -3e-4
    m._k = lambda h, q: lam * q
    q0 = np.ones((12, 12))
    _, q1 = m.one_step(np.zeros((12, 12)), q0, np.zeros((12, 12)), np.zeros((12, 12)))
    np.testing.assert_allclose(q1, poly(lam * m.dt), rtol=1e-14)


# ------------------------------------------------------------- linearisation
@pytest.mark.parametrize("scheme", ["Euler", "rk2", "rk3"])
@pytest.mark.parametrize("opts", [{}, {"Kdiffus": 30.0}, {"land": True},
                                  {"beta": True, "bathy": True, "mdt": True}])
def test_adjoint_dot_product_and_tangent(scheme, opts):
    rng = np.random.default_rng(5)
    ny, nx = 22, 26
    kw = {}
    ssh = None
    if opts.get("land"):
        ssh = np.zeros((ny, nx))
        ssh[9, 11] = np.nan
        kw["SSH"] = ssh
    if "Kdiffus" in opts:
        kw["Kdiffus"] = opts["Kdiffus"]
    f = F0
    if opts.get("beta"):
        f = F0 + 1.6e-11 * 7e3 * (np.arange(ny)[:, None] - ny / 2) * np.ones((1, nx))
    if opts.get("bathy"):
        kw["bathymetry_PV_term"] = 0.1 * rng.standard_normal((ny, nx))
    if opts.get("mdt"):
        kw["mdt"] = 0.05 * rng.standard_normal((ny, nx))
    m = QgmOracle(dx=10e3, dy=7e3, dt=600.0, c=2.5, f=f, time_scheme=scheme, shape=(ny, nx), **kw)
    h0 = 0.05 * rng.standard_normal((ny, nx))
    if ssh is not None:
        h0[np.isnan(ssh)] = np.nan
    hb = 0.01 * rng.standard_normal((ny, nx))
    dh = 0.05 * rng.standard_normal((ny, nx))
    y = rng.standard_normal((ny, nx))
    n = 4
    tl, h1 = m.step_tgl(dh, h0, hb, n)
    ad = m.step_adj(y, h0, hb, n)
    ok = ~np.isnan(h1)
    p1, p2 = np.sum(tl[ok] * y[ok]), np.nansum(dh * ad)
    assert abs(p1 / p2 - 1) < 1e-12
    # centred finite differences of the forward agree with the tangent
    eps = 1e-6
    fd = (m.step(h0 + eps * dh, hb, n) - m.step(h0 - eps * dh, hb, n)) / (2 * eps)
    assert np.linalg.norm((fd - tl)[ok]) / np.linalg.norm(tl[ok]) < 1e-5


def test_where_subgradient_at_zero_velocity():
    """At h == 0 every uT is exactly 0, and JAX's where-rule gives both up and um
    derivative 1; the tangent is then adv(duT-split, q) with both branches."""
    m = model(20, 20, "Euler")
    rng = np.random.default_rng(3)
    h0 = np.zeros((20, 20))
    dh = 0.05 * rng.standard_normal((20, 20))
    hb = np.zeros((20, 20))
    q = m.h2pv(h0, hb)  # zero
    u, v = m.h2uv(h0)
    du, dv = m.h2uv_tl(dh, h0)
    dq = m.h2pv_tl(dh, h0)
    qq = q + 1e-6 * rng.standard_normal(q.shape)
    dr = m.rhs_tl(du, dv, dq, u, v, qq)
    duT = 0.5 * (du[1:-1, 1:-1] + du[1:-1, 2:])
    dvT = 0.5 * (dv[1:-1, 1:-1] + dv[2:, 1:-1])
    expect = m.adv(duT, dvT, duT, dvT, qq)
    expect[m.ind12] = 0
    expect[m.ind0] = 0
    np.testing.assert_allclose(dr, expect, atol=1e-25)
