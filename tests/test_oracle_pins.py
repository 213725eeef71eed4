"""Closed-form pins for the optional branches of the oracle's ``rhs`` (beta
plane, PV diffusion, bathymetry; jqgm.py:350-369), for the flux-control
increment of ``ModelOracle`` (mod.py:911-913, 1058-1068), and a brute-force
point-by-point evaluation of the stencils on tiny grids.

Each expected value is fixed by the mathematics, not by the oracle's code:
centred differences are exact for quadratics (the beta plane), the 5-point
Laplacian for cubics (diffusion), the 3rd-order upwind scheme for cubics
(bathymetry), and a model at rest stays at rest (the flux increment).  Grids
are rectangular with dx != dy, and the polynomials have different x and y
coefficients, so a swapped spacing, a shifted row or a dropped term fails.
"""
import numpy as np
import pytest

from oracle.fourdvar import ModelOracle
from oracle.qg import QgmOracle

G, F0 = 9.81, 1e-4
DX, DY = 10e3, 7e3


def grid(ny, nx):
    y = DY * np.arange(ny)[:, None] * np.ones((1, nx))
    x = DX * np.arange(nx)[None, :] * np.ones((ny, 1))
    return y, x


def oracle(ny=22, nx=26, **kw):
    return QgmOracle(dx=DX, dy=DY, dt=600.0, c=2.5, time_scheme="rk2", shape=(ny, nx), **kw)


# --------------------------------------------------------------- beta plane
def test_rhs_beta_plane_closed_form():
    """f = f0 + b0 y + g2 y^2: the centred beta is b0 + 2 g2 y_i exactly, and the
    beta term at row i uses the mean of v at rows i and i+1."""
    ny, nx = 22, 26
    y, x = grid(ny, nx)
    b0, g2 = 1.6e-11, 3e-17
    f = F0 + b0 * y + g2 * y ** 2
    m = oracle(ny, nx, f=f)
    V0, V1 = 0.3, -0.02
    v = V0 + V1 * np.arange(ny)[:, None] * np.ones((1, nx))
    u = 0.7 * np.ones((ny, nx))  # irrelevant: q = 0 makes the advection vanish
    r = m.rhs(u, v, np.zeros((ny, nx)))
    i = np.arange(ny)[:, None] * np.ones((1, nx))
    expect = -(b0 + 2 * g2 * y) * (V0 + V1 * (i + 0.5))
    act = m.mask == 3
    np.testing.assert_allclose(r[act], expect[act], rtol=1e-9)
    assert np.all(r[~act] == 0)


def test_rhs_beta_sign_with_uniform_meridional_flow():
    """Linear f(y) and uniform v: rhs = -beta0 * v on the active points."""
    ny, nx = 20, 24
    y, _ = grid(ny, nx)
    b0 = 2e-11
    m = oracle(ny, nx, f=F0 + b0 * y)
    V = -0.4
    r = m.rhs(np.zeros((ny, nx)), V * np.ones((ny, nx)), np.zeros((ny, nx)))
    act = m.mask == 3
    np.testing.assert_allclose(r[act], -b0 * V, rtol=1e-9)


# ------------------------------------------------------------ PV diffusion
def test_rhs_diffusion_closed_form():
    """No wind, cubic q: rhs = K (d2q/dx2 + d2q/dy2) exactly (5-point Laplacian)."""
    ny, nx = 22, 26
    y, x = grid(ny, nx)
    K = 40.0
    m = oracle(ny, nx, f=F0, Kdiffus=K)
    a, b, c, d = 3e-15, -7e-15, 2e-10, 5e-21
    q = a * x ** 2 + b * y ** 2 + c * x * y * 1e-5 + d * x ** 3 + 1e-6 * y
    lap = 2 * a + 2 * b + 6 * d * x
    zero = np.zeros((ny, nx))
    r = m.rhs(zero, zero, q)
    act = m.mask == 3
    np.testing.assert_allclose(r[act], (K * lap)[act], rtol=1e-7)
    assert np.all(r[~act] == 0)


# --------------------------------------------------------------- bathymetry
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_rhs_bathymetry_closed_form(sign):
    """q = 0, constant winds, cubic bathymetry term b: rhs = -f0 (U db/dx + V db/dy)
    (the reference advects f * bathymetry_PV_term with the same upwind winds)."""
    ny, nx = 22, 26
    y, x = grid(ny, nx)
    bathy = 2e-13 * x ** 3 - 1e-8 * x ** 2 + 5e-13 * y ** 3 + 4e-4 * y
    dbdx = 6e-13 * x ** 2 - 2e-8 * x
    dbdy = 1.5e-12 * y ** 2 + 4e-4
    m = oracle(ny, nx, f=F0, bathymetry_PV_term=bathy)
    U, V = 0.35 * sign, -0.2 * sign
    r = m.rhs(U * np.ones((ny, nx)), V * np.ones((ny, nx)), np.zeros((ny, nx)))
    expect = -F0 * (U * dbdx + V * dbdy)
    act = m.mask == 3
    np.testing.assert_allclose(r[act], expect[act], rtol=1e-8)


def test_rhs_all_branches_superpose():
    """All branches together: advection of a cubic q and of f_mean * bathymetry,
    the beta term and diffusion add up to the sum of their closed forms; the
    bathymetry term uses the NaN-mean of f (jqgm.py:141, 353-354)."""
    ny, nx = 24, 28
    y, x = grid(ny, nx)
    b0 = 1.6e-11
    f = F0 + b0 * (y - y.mean())
    fbar = np.nanmean(f)
    K = 25.0
    bathy = 1e-13 * x ** 3 + 2e-13 * y ** 3
    q = 1e-15 * x ** 3 - 2e-10 * x ** 2 + 3e-16 * y ** 3 + 1e-9 * y
    m = oracle(ny, nx, f=f, Kdiffus=K, bathymetry_PV_term=bathy)
    U, V = 0.25, 0.15
    r = m.rhs(U * np.ones((ny, nx)), V * np.ones((ny, nx)), q)
    dqdx = 3e-15 * x ** 2 - 4e-10 * x
    dqdy = 9e-16 * y ** 2 + 1e-9
    lapq = 6e-15 * x - 4e-10 + 18e-16 * y
    expect = (-(U * dqdx + V * dqdy) - fbar * (U * 3e-13 * x ** 2 + V * 6e-13 * y ** 2)
              - b0 * V + K * lapq)
    act = m.mask == 3
    np.testing.assert_allclose(r[act], expect[act], rtol=1e-7)


# ----------------------------------------------------------- flux control
@pytest.mark.parametrize("n", [1, 7])
def test_flux_increment_from_rest(n):
    """From rest with a zero boundary field the QG segment stays exactly zero,
    so one segment with flux F gives h1 = n dt / 86400 * F: the increment is
    added after the segment (not advected) and scales with the segment length."""
    ny, nx = 18, 22
    qg = oracle(ny, nx, f=F0)
    M = ModelOracle(qg)
    F = np.random.default_rng(4).standard_normal((ny, nx))
    h1 = M.step(np.zeros((ny, nx)), F, n)
    np.testing.assert_allclose(h1, n * qg.dt / 86400.0 * F, rtol=1e-15, atol=0)


def test_flux_increment_is_additive_after_the_segment():
    ny, nx = 18, 22
    qg = oracle(ny, nx, f=F0)
    M = ModelOracle(qg)
    rng = np.random.default_rng(5)
    h0 = 0.05 * rng.standard_normal((ny, nx))
    F = rng.standard_normal((ny, nx))
    n = 5
    d = M.step(h0, F, n) - M.step(h0, np.zeros((ny, nx)), n)
    np.testing.assert_allclose(d, n * qg.dt / 86400.0 * F, rtol=1e-12, atol=1e-15)


def test_flux_adjoint_uses_pre_propagation_adjoint():
    """adFssh += n dt/86400 * adX1 with adX1 the adjoint BEFORE the segment's
    adjoint (mod.py:1058-1068); the state adjoint is the QG segment adjoint.
    The pair is the transpose of (h0, F) -> h1: checked by the dot product."""
    ny, nx = 18, 22
    qg = oracle(ny, nx, f=F0)
    M = ModelOracle(qg)
    rng = np.random.default_rng(6)
    h0 = 0.05 * rng.standard_normal((ny, nx))
    y = rng.standard_normal((ny, nx))
    n = 4
    scale = n * qg.dt / 86400.0
    adh0, adF = M.step_adj(y.copy(), np.zeros((ny, nx)), h0, n)
    np.testing.assert_allclose(adF, scale * y, rtol=1e-15)
    assert np.linalg.norm(adh0 - y) > 1e-3 * np.linalg.norm(y)  # the segment changes it
    dh, dF = rng.standard_normal((2, ny, nx))
    tl, _ = qg.step_tgl(dh, h0, np.zeros((ny, nx)), n)
    lhs = np.sum((tl + scale * dF) * y)
    rhs_ = np.sum(dh * adh0) + np.sum(dF * adF)
    assert abs(lhs / rhs_ - 1) < 1e-12


# ------------------------------------------------- brute force, point by point
def brute_h2uv(h, g, f, dx, dy):
    ny, nx = h.shape
    u = np.zeros((ny, nx))
    v = np.zeros((ny, nx))
    for i in range(ny):
        for j in range(nx):
            if 1 <= i <= ny - 2 and 1 <= j <= nx - 1:
                u[i, j] = -g / f * (h[i + 1, j - 1] + h[i + 1, j] - h[i - 1, j] - h[i - 1, j - 1]) / (4 * dy)
            if 1 <= i <= ny - 1 and 1 <= j <= nx - 2:
                v[i, j] = g / f * (h[i, j + 1] + h[i - 1, j + 1] - h[i - 1, j - 1] - h[i, j - 1]) / (4 * dx)
    return u, v


def brute_rhs(m, u, v, q, beta=None, K=None, fb=None):
    ny, nx = q.shape
    dx, dy = m.dx, m.dy
    r = np.zeros((ny, nx))

    def d_up(a, i, j, ax):
        # 3rd-order upwind pieces at (i, j) along axis ax (0: y, 1: x)
        def at(k):
            return a[i + k, j] if ax == 0 else a[i, j + k]
        p = 2 * at(1) + 3 * at(0) - 6 * at(-1) + at(-2)
        mm = at(2) - 6 * at(1) + 3 * at(0) + 2 * at(-1)
        return p, mm

    for i in range(2, ny - 2):
        for j in range(2, nx - 2):
            uT = 0.5 * (u[i, j] + u[i, j + 1])
            vT = 0.5 * (v[i, j] + v[i + 1, j])
            up, um = max(uT, 0.0), min(uT, 0.0)
            vp, vm = max(vT, 0.0), min(vT, 0.0)
            val = 0.0
            for fld in ([q] if fb is None else [q, fb]):
                px, mx = d_up(fld, i, j, 1)
                py, my = d_up(fld, i, j, 0)
                val += -up * px / (6 * dx) + um * mx / (6 * dx) - vp * py / (6 * dy) + vm * my / (6 * dy)
            if beta is not None:
                val -= beta[i] * (v[i, j] + v[i + 1, j]) / 2
            if K is not None:
                val += K / dx ** 2 * (q[i, j + 1] + q[i, j - 1] - 2 * q[i, j]) + \
                    K / dy ** 2 * (q[i + 1, j] + q[i - 1, j] - 2 * q[i, j])
            r[i, j] = val
    r[m.mask != 3] = 0
    return r


@pytest.mark.parametrize("shape", [(7, 9), (8, 10)])
def test_stencils_match_point_by_point_loops(shape):
    ny, nx = shape
    rng = np.random.default_rng(7)
    y, _ = grid(ny, nx)
    fcol = F0 + 1.6e-11 * y + 1e-17 * y ** 2
    beta_rows = np.zeros(ny)
    for i in range(1, ny - 1):
        beta_rows[i] = (fcol[i + 1, 0] - fcol[i - 1, 0]) / (2 * DY)
    bathy = 0.1 * rng.standard_normal((ny, nx))
    K = 30.0
    m = oracle(ny, nx, f=fcol, Kdiffus=K, bathymetry_PV_term=bathy)
    h = 0.1 * rng.standard_normal((ny, nx))
    u, v = m.h2uv(h)
    bu, bv = brute_h2uv(h, G, m.f, DX, DY)
    np.testing.assert_allclose(u, bu, rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(v, bv, rtol=1e-13, atol=1e-16)
    hb = 0.01 * rng.standard_normal((ny, nx))
    q = m.h2pv(h, hb)
    bq = np.zeros((ny, nx))
    for i in range(1, ny - 1):
        for j in range(1, nx - 1):
            bq[i, j] = G / m.f * ((h[i + 1, j] + h[i - 1, j] - 2 * h[i, j]) / DY ** 2 +
                                  (h[i, j + 1] + h[i, j - 1] - 2 * h[i, j]) / DX ** 2) - \
                G * m.f / m.c ** 2 * h[i, j]
            if m.mask[i, j] in (1, 2):
                bq[i, j] = -G * m.f / m.c ** 2 * hb[i, j]
    bq[m.mask == 1] = -G * m.f / m.c ** 2 * hb[m.mask == 1]
    np.testing.assert_allclose(q, bq, rtol=1e-13, atol=1e-20)
    qq = 1e-6 * rng.standard_normal((ny, nx))
    r = m.rhs(u, v, qq)
    br = brute_rhs(m, u, v, qq, beta=beta_rows, K=K, fb=m.f * bathy)
    np.testing.assert_allclose(r, br, rtol=1e-12, atol=1e-24)
