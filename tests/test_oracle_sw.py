"""Pins of the shallow-water oracle (oracle/sw.py) by what the mathematics and
the reference's definitions fix, independently of the oracle's own code.

* Masks: the stencil classes of a rectangular basin derived by hand from the
  window definitions of masks.py:95-153 (2-point next to a wall, 4-point one
  further in, 6-point where three cells lie on each side; the vorticity flux
  along y on the u grid marks its 4-point faces for the 6-point
  reconstruction too, sw.py:285-293), and for a basin with land, the same
  classes computed by counting ocean cells in the windows (a different
  formulation of the avg_pool2d thresholds).
* Reconstructions: both WENO-Z reconstructions are exact for linear data in
  either direction, the 6-point one for cell averages of a quadratic.
* RHS closed forms: the rest state stays at rest; a uniform flow on an
  f-plane gets the Coriolis acceleration (f v, -f u) away from the walls; a
  thickness ramp in one layer gives the pressure gradient of the layer
  coupling (-g'_0 a in every layer for layer 0, -(g'_0 + g'_1) a below layer
  0 for a ramp in layer 1); a uniform flow across a thickness ramp gives
  dh/dt = -u a in the inner cells and zero in the first and last cell (the
  edge-padded divergence of finite_diff.py:42-52).
* Vorticity: -a / b at interior points for the linear shears u = a y,
  v = b x; -2 (1 - slip) U0 / dy at the bottom wall (+ at the top) for a
  uniform flow, for no-slip, partial and free slip.
* The mirrors x -> -x (u -> -u, f -> -f) and y -> -y (v -> -v, f(y) ->
  -f(-y)) map a run onto the mirrored run.
* The complex-step tangent equals central finite differences; one RK3-SSP
  step of a linear RHS is the cubic Taylor polynomial of the exponential.

Parity unpinned: the vorticity flux along y blending the 6-point
reconstruction at its 4-point faces (the reference's mask choice,
sw.py:285-293) changes results only next to land, where no closed form
applies; it is restated from the source and the GPU model follows it
(tests/test_gpu_sw.py, basins with islands).
"""
import numpy as np
import pytest

from oracle.sw import SwOracle, sw_masks, wenoz4_left, wenoz6_left

NONE, C2, C4, C6, C46 = 0, 2, 4, 6, 10


def classes(m, pre):
    c = np.zeros(m[pre + "_eq2"].shape, dtype=int)
    c[m[pre + "_eq2"] > 0] = C2
    c[m[pre + "_eq4"] > 0] = C4
    six = m[pre + ("_gt4" if pre == "u_sten_wy" else "_gt6")]
    c[(six > 0) & (c == C4)] = C46
    c[(six > 0) & (c == NONE)] = C6
    return c


def test_masks_rectangle_known_answer():
    nx, ny = 8, 7
    m = sw_masks(np.ones((nx, ny)))
    u = np.ones((nx + 1, ny)); u[[0, nx], :] = 0
    v = np.ones((nx, ny + 1)); v[:, [0, ny]] = 0
    np.testing.assert_array_equal(m["u"], u)
    np.testing.assert_array_equal(m["v"], v)
    along_x = [NONE, C2, C4, C6, C6, C6, C4, C2, NONE]            # u points i = 0..8
    along_y = [NONE, C2, C4, C6, C6, C4, C2, NONE]                # v points j = 0..7
    np.testing.assert_array_equal(classes(m, "u_sten_hx"), np.repeat(np.array(along_x)[:, None], ny, 1))
    np.testing.assert_array_equal(classes(m, "v_sten_hy"), np.repeat(np.array(along_y)[None, :], nx, 0))
    # vorticity stencils over the nx+1 / ny+1 w points (the walls' w points count)
    wy = [C2, C46, C6, C6, C6, C46, C2]                          # u points j = 0..6
    exp = np.repeat(np.array(wy)[None, :], nx + 1, 0); exp[[0, nx], :] = NONE
    np.testing.assert_array_equal(classes(m, "u_sten_wy"), exp)
    wx = [C2, C4, C6, C6, C6, C6, C4, C2]                        # v points i = 0..7
    exp = np.repeat(np.array(wx)[:, None], ny + 1, 1); exp[:, [0, ny]] = NONE
    np.testing.assert_array_equal(classes(m, "v_sten_wx"), exp)
    # vorticity boundary classes: walls without the corners; corners and interior valid
    vb = np.zeros((nx + 1, ny + 1)); vb[[0, nx], 1:ny] = 1
    hb = np.zeros((nx + 1, ny + 1)); hb[1:nx, [0, ny]] = 1
    np.testing.assert_array_equal(m["w_vertical_bound"], vb)
    np.testing.assert_array_equal(m["w_horizontal_bound"], hb)
    np.testing.assert_array_equal(m["w_cornerout_bound"], 0 * vb)
    np.testing.assert_array_equal(m["w_valid"], 1 - vb - hb)


def window_all_ocean(h, axis, lo, hi):
    """1 where every cell at offsets lo..hi (inclusive) along axis is ocean."""
    n = h.shape[axis]
    out = np.ones([s + (1 if d == axis else 0) for d, s in enumerate(h.shape)], dtype=bool)
    idx = np.arange(out.shape[axis])
    for off in range(lo, hi + 1):
        c = idx + off
        ok = (c >= 0) & (c < n)
        vals = np.take(h, np.clip(c, 0, n - 1), axis=axis) > 0
        shape = [1] * h.ndim; shape[axis] = -1
        out &= vals & ok.reshape(shape)
    return out


def test_masks_with_land_by_counting():
    h = np.ones((10, 9)); h[3:6, 2:4] = 0; h[0, 0] = 0; h[9, 5:] = 0; h[6, 7] = 0
    m = sw_masks(h)
    # u / v: both neighbouring cells ocean
    np.testing.assert_array_equal(m["u"] > 0, window_all_ocean(h, 0, -1, 0))
    np.testing.assert_array_equal(m["v"] > 0, window_all_ocean(h, 1, -1, 0))
    # h stencils along x on u points: 2 where the 4-cell window is not all
    # ocean, 4 where the 6-cell one is not, else 6
    two, four, six = (window_all_ocean(h, 0, -k, k - 1) for k in (1, 2, 3))
    exp = np.where(two & ~four, C2, np.where(four & ~six, C4, np.where(six, C6, NONE)))
    np.testing.assert_array_equal(classes(m, "u_sten_hx"), exp)
    two, four, six = (window_all_ocean(h, 1, -k, k - 1) for k in (1, 2, 3))
    exp = np.where(two & ~four, C2, np.where(four & ~six, C4, np.where(six, C6, NONE)))
    np.testing.assert_array_equal(classes(m, "v_sten_hy"), exp)


@pytest.mark.parametrize("a,b", [(0.3, 1.7), (-2.0, 0.25)])
def test_weno_exact_for_linear_data(a, b):
    q = a + b * np.arange(6.0)
    # face between points 2 and 3 at position 2.5: left-biased from the left
    # (q[1..3] / q[0..4]) and from the right (reversed stencils)
    assert wenoz4_left(q[1], q[2], q[3], q[4]) == pytest.approx(a + 2.5 * b, rel=1e-13)
    assert wenoz4_left(q[4], q[3], q[2], q[1]) == pytest.approx(a + 2.5 * b, rel=1e-13)
    assert wenoz6_left(*q) == pytest.approx(a + 2.5 * b, rel=1e-13)
    assert wenoz6_left(*q[::-1]) == pytest.approx(a + 2.5 * b, rel=1e-13)


def test_weno6_exact_for_cell_averages_of_a_quadratic():
    c0, c1, c2 = 0.7, -0.3, 0.05
    P = lambda x: c0 * x + c1 * x ** 2 / 2 + c2 * x ** 3 / 3  # primitive of the quadratic
    avg = np.array([P(k + 0.5) - P(k - 0.5) for k in range(6)])
    x = 2.5
    assert wenoz6_left(*avg) == pytest.approx(c0 + c1 * x + c2 * x ** 2, rel=1e-12)
    assert wenoz6_left(*avg[::-1]) == pytest.approx(c0 + c1 * x + c2 * x ** 2, rel=1e-12)


def basin(nx=16, ny=14, nl=3, beta=1.6e-11, f0=1e-4, mask=None, slip=0.0, drag=0.0, dt=120.0):
    dx = dy = 8e3
    y = (np.arange(ny + 1) - ny / 2) * dy
    f = np.broadcast_to(f0 + beta * y, (nx + 1, ny + 1)).copy()
    H = np.array([500., 1000., 3000.][:nl]); gp = np.array([0.02, 0.02, 0.01][:nl])
    return SwOracle(nx, ny, dx, dy, H, gp, f, dt, mask=mask, slip_coef=slip, bottom_drag_coef=drag)


def test_rest_state_stays_at_rest():
    h = np.ones((16, 14)); h[5:8, 4:7] = 0
    orc = basin(mask=h)
    z = [np.zeros((3, 17, 14)), np.zeros((3, 16, 15)), np.zeros((3, 16, 14))]
    out = orc.step(*z, nstep=3)
    for o in out:
        assert np.all(o == 0)


@pytest.mark.parametrize("comp", ["u", "v"])
def test_uniform_flow_coriolis(comp):
    orc = basin(beta=0.0, f0=1e-4)
    nl, nx, ny = 3, 16, 14
    U0 = 0.2
    u = np.zeros((nl, nx + 1, ny)); v = np.zeros((nl, nx, ny + 1)); h = np.zeros((nl, nx, ny))
    if comp == "u":
        u[:] = U0
    else:
        v[:] = U0
    du, dv, dh = orc.rhs(*orc.to_scaled(u, v, h))
    du, dv, dh = du / orc.dx, dv / orc.dy, dh / orc.area   # physical tendencies
    inner = (slice(None), slice(4, -4), slice(4, -4))
    if comp == "u":   # dv/dt = -f u, du/dt = 0
        np.testing.assert_allclose(dv[inner], -1e-4 * U0, rtol=1e-12)
        np.testing.assert_allclose(du[inner], 0.0, atol=1e-18)
    else:             # du/dt = f v
        np.testing.assert_allclose(du[inner], 1e-4 * U0, rtol=1e-12)
        np.testing.assert_allclose(dv[inner], 0.0, atol=1e-18)
    np.testing.assert_allclose(dh, 0.0, atol=1e-18)


@pytest.mark.parametrize("layer", [0, 1])
def test_pressure_gradient_couples_layers(layer):
    orc = basin(beta=0.0, f0=0.0)
    nl, nx, ny = 3, 16, 14
    a = 1e-5  # m per m
    x = (np.arange(nx) + 0.5) * orc.dx
    h = np.zeros((nl, nx, ny)); h[layer] = a * x[:, None]
    du, dv, dh = orc.rhs(np.zeros((nl, nx + 1, ny)), np.zeros((nl, nx, ny + 1)), orc.to_scaled(0, 0, h)[2])
    du = du / orc.dx
    gp = orc.gp
    expect = [-gp[0] * a] * nl if layer == 0 else [-gp[0] * a, -(gp[0] + gp[1]) * a, -(gp[0] + gp[1]) * a]
    for l in range(nl):
        np.testing.assert_allclose(du[l, 1:nx, :], expect[l], rtol=1e-10)
        assert np.all(du[l, [0, nx], :] == 0)
    np.testing.assert_allclose(dv / orc.dy, 0.0, atol=1e-18)


def test_advection_of_a_ramp_and_the_boundary_cells():
    orc = basin(nl=1, beta=0.0, f0=0.0)
    nx, ny = 16, 14
    U0, a = 0.1, 1e-5
    x = (np.arange(nx) + 0.5) * orc.dx
    u = np.full((1, nx + 1, ny), U0)
    h = (a * x)[None, :, None] * np.ones((1, nx, ny))
    _, _, dh = orc.rhs(*orc.to_scaled(u, np.zeros((1, nx, ny + 1)), h))
    dh = dh / orc.area
    np.testing.assert_allclose(dh[0, 1:nx - 1, :], -U0 * a, rtol=1e-6)
    assert np.all(dh[0, [0, nx - 1], :] == 0)


def test_mirror_symmetry():
    rng = np.random.default_rng(4)
    nl, nx, ny = 2, 16, 14
    o1 = basin(nl=nl, beta=1.6e-11, f0=1e-4)
    o2 = basin(nl=nl, beta=-1.6e-11, f0=-1e-4)
    m = o1.m
    u = 0.1 * rng.standard_normal((nl, nx + 1, ny)) * m["u"]
    v = 0.1 * rng.standard_normal((nl, nx, ny + 1)) * m["v"]
    h = 0.5 * rng.standard_normal((nl, nx, ny))
    a = o1.step(u, v, h, nstep=4)
    b = o2.step(-u[:, ::-1], v[:, ::-1], h[:, ::-1], nstep=4)
    np.testing.assert_allclose(-b[0][:, ::-1], a[0], rtol=0, atol=1e-12 * np.abs(a[0]).max())
    np.testing.assert_allclose(b[1][:, ::-1], a[1], rtol=0, atol=1e-12 * np.abs(a[1]).max())
    np.testing.assert_allclose(b[2][:, ::-1], a[2], rtol=0, atol=1e-12 * np.abs(a[2]).max())


def test_complex_step_tangent_equals_finite_differences():
    rng = np.random.default_rng(1)
    h0 = np.ones((16, 14)); h0[6:9, 3:5] = 0
    orc = basin(mask=h0, slip=0.5, drag=1e-6)
    m = orc.m
    x = [0.1 * rng.standard_normal((3, 17, 14)) * m["u"], 0.1 * rng.standard_normal((3, 16, 15)) * m["v"],
         0.5 * rng.standard_normal((3, 16, 14)) * m["h"]]
    d = [rng.standard_normal(a.shape) * mk for a, mk in zip(x, (m["u"], m["v"], m["h"]))]
    tl = orc.step_tgl(*x, *d, nstep=2)
    eps = 1e-6
    p = orc.step(*(a + eps * b for a, b in zip(x, d)), nstep=2)
    q = orc.step(*(a - eps * b for a, b in zip(x, d)), nstep=2)
    for t, a, b in zip(tl, p, q):
        fd = (a - b) / (2 * eps)
        assert np.linalg.norm(t - fd) <= 1e-6 * np.linalg.norm(t)


def test_rk3_is_the_cubic_taylor_polynomial():
    orc = basin(nl=1)
    lam = -0.3 + 0.8j
    z = lam * orc.dt / 1000.0
    orc.rhs = lambda u, v, h: tuple(lam / 1000.0 * a for a in (u, v, h))
    x = tuple(np.ones(3, dtype=complex) for _ in range(3))
    out = orc.rk3(*x)
    assert out[0][0] == pytest.approx(1 + z + z ** 2 / 2 + z ** 3 / 6, rel=1e-14)


@pytest.mark.parametrize("slip", [0.0, 0.5, 1.0])
def test_vorticity_closed_forms(slip):
    orc = basin(nl=1, slip=slip)
    nx, ny = 16, 14
    a, b, U0 = 2e-6, -3e-6, 0.2
    y_u = (np.arange(ny) + 0.5) * orc.dy
    x_v = (np.arange(nx) + 0.5) * orc.dx
    # linear shears: omega = -du/dy, dv/dx at interior w points
    u = np.zeros((1, nx + 1, ny)); u[0, 1:nx, :] = a * y_u
    w = orc.omega(*orc.to_scaled(u, np.zeros((1, nx, ny + 1)), 0)[:2]) / orc.area
    np.testing.assert_allclose(w[0, 1:nx, 1:ny], -a, rtol=1e-10)
    v = np.zeros((1, nx, ny + 1)); v[0, :, 1:ny] = b * x_v[:, None]
    w = orc.omega(*orc.to_scaled(np.zeros((1, nx + 1, ny)), v, 0)[:2]) / orc.area
    np.testing.assert_allclose(w[0, 1:nx, 1:ny], b, rtol=1e-10)
    # wall vorticity of a uniform flow: -2 (1 - slip) U0 / dy on the bottom
    # wall, +2 (1 - slip) U0 / dy on the top one, zero in the interior
    u = np.full((1, nx + 1, ny), U0)
    w = orc.omega(*orc.to_scaled(u, np.zeros((1, nx, ny + 1)), 0)[:2]) / orc.area
    np.testing.assert_allclose(w[0, 1:nx, 0], -2 * (1 - slip) * U0 / orc.dy, rtol=1e-12, atol=1e-20)
    np.testing.assert_allclose(w[0, 1:nx, ny], 2 * (1 - slip) * U0 / orc.dy, rtol=1e-12, atol=1e-20)
    np.testing.assert_allclose(w[0, 1:nx, 1:ny], 0.0, atol=1e-20)


def test_mirror_symmetry_y():
    rng = np.random.default_rng(5)
    nl, nx, ny = 2, 16, 14
    o1 = basin(nl=nl, beta=1.6e-11, f0=1e-4)
    o2 = basin(nl=nl, beta=1.6e-11, f0=-1e-4)   # f2(y) = -f1(-y)
    m = o1.m
    u = 0.1 * rng.standard_normal((nl, nx + 1, ny)) * m["u"]
    v = 0.1 * rng.standard_normal((nl, nx, ny + 1)) * m["v"]
    h = 0.5 * rng.standard_normal((nl, nx, ny))
    a = o1.step(u, v, h, nstep=4)
    b = o2.step(u[:, :, ::-1], -v[:, :, ::-1], h[:, :, ::-1], nstep=4)
    np.testing.assert_allclose(b[0][:, :, ::-1], a[0], rtol=0, atol=1e-12 * np.abs(a[0]).max())
    np.testing.assert_allclose(-b[1][:, :, ::-1], a[1], rtol=0, atol=1e-12 * np.abs(a[1]).max())
    np.testing.assert_allclose(b[2][:, :, ::-1], a[2], rtol=0, atol=1e-12 * np.abs(a[2]).max())


def _basin(nx=12, ny=10, nl=3, fplane=False):
    dx = dy = 8e3
    y = (np.arange(ny + 1) - ny / 2) * dy
    f = np.broadcast_to(1e-4 + (0.0 if fplane else 1.6e-11) * y, (nx + 1, ny + 1)).copy()
    H = np.array([500.0, 1000.0, 3000.0][:nl])
    gp = np.array([0.02, 0.02, 0.01][:nl])
    return SwOracle(nx, ny, dx, dy, H, gp, f, 120.0)


def test_forcing_uniform_thickness_is_exact():
    """A uniform thickness forcing of layer 0 from rest has no pressure
    gradient and no flux: after n steps h_0 = n dt F and nothing else moves
    (the forcing is x += dt F after every RK3 step, sw.py:875-877)."""
    o = _basin()
    nl, nx, ny = o.nl, o.nx, o.ny
    z = (np.zeros((nl, nx + 1, ny)), np.zeros((nl, nx, ny + 1)), np.zeros((nl, nx, ny)))
    Fh = np.zeros((nl, nx, ny)); Fh[0] = 3e-6
    for n in (1, 4):
        u, v, h = o.step(*z, nstep=n, F=(np.zeros_like(z[0]), np.zeros_like(z[1]), Fh))
        assert np.all(u == 0) and np.all(v == 0) and np.all(h[1:] == 0)
        np.testing.assert_allclose(h[0], n * o.dt * 3e-6, rtol=1e-13)


def test_forcing_one_step_from_rest_is_dt_F_masked():
    """From rest the RK3 stages have zero tendencies, so one step gives
    dt F on the ocean points of each grid (u and v zero on the walls)."""
    o = _basin(nl=2)
    rng = np.random.default_rng(4)
    nl, nx, ny = o.nl, o.nx, o.ny
    F = (rng.standard_normal((nl, nx + 1, ny)), rng.standard_normal((nl, nx, ny + 1)),
         rng.standard_normal((nl, nx, ny)))
    z = tuple(np.zeros_like(a) for a in F)
    out = o.step(*z, nstep=1, F=F)
    for got, Fi, m in zip(out, F, (o.m["u"], o.m["v"], o.m["h"])):
        np.testing.assert_allclose(got, o.dt * Fi * m, rtol=1e-13, atol=1e-30)


def test_forcing_tangent_equals_finite_differences():
    o = _basin(nl=2)
    rng = np.random.default_rng(8)
    nl, nx, ny = o.nl, o.nx, o.ny
    m = o.m
    x = [0.05 * rng.standard_normal((nl, nx + 1, ny)) * m["u"],
         0.05 * rng.standard_normal((nl, nx, ny + 1)) * m["v"], 0.3 * rng.standard_normal((nl, nx, ny)) * m["h"]]
    F = [1e-6 * rng.standard_normal(a.shape) for a in x]
    dF = [1e-6 * rng.standard_normal(a.shape) for a in x]
    zero = [np.zeros_like(a) for a in x]
    tl = o.step_tgl(*x, *zero, nstep=3, F=F, dF=dF)
    eps = 1e-3
    p = o.step(*x, nstep=3, F=[a + eps * d for a, d in zip(F, dF)])
    q = o.step(*x, nstep=3, F=[a - eps * d for a, d in zip(F, dF)])
    for t, a, b in zip(tl, p, q):
        fd = (a - b) / (2 * eps)
        assert np.linalg.norm(t - fd) <= 1e-6 * np.linalg.norm(fd)


@pytest.mark.parametrize("hw2d", [False, True])
def test_wind_tendency_at_rest(hw2d):
    """At rest the tendency is the wind term alone (add_wind_forcing,
    sw.py:567-642): tau / (rho_water H_ref) times dx (u) or dy (v) on the top
    layer's interior ocean points, H_ref the scalar h_wind clamped at h_min or
    the mean of the two h-cells around the face; nothing below the top layer
    and nothing in h.  dx != dy, so a swapped metric fails."""
    nx, ny, nl, dx, dy = 12, 10, 3, 6e3, 9e3
    f = np.full((nx + 1, ny + 1), 1e-4)
    mask = np.ones((nx, ny)); mask[4:6, 3:5] = 0
    rng = np.random.default_rng(5)
    taux = 0.1 + 0.05 * rng.standard_normal((nx - 1, ny))
    tauy = 0.07
    hw = (40.0 + 30.0 * rng.random((nx, ny))) if hw2d else 0.05  # the scalar below h_min = 0.1
    orc = SwOracle(nx, ny, dx, dy, np.array([500., 1000., 3000.]), np.array([0.02, 0.02, 0.01]), f, 120.0,
                   mask=mask, taux=taux, tauy=tauy, h_wind=hw, rho_water=1000.0)
    du, dv, dh = orc.rhs(np.zeros((nl, nx + 1, ny)), np.zeros((nl, nx, ny + 1)), np.zeros((nl, nx, ny)))
    Hu = 0.5 * (hw[:-1, :] + hw[1:, :]) if hw2d else 0.1
    Hv = 0.5 * (hw[:, :-1] + hw[:, 1:]) if hw2d else 0.1
    mu, mv = orc.m["u"], orc.m["v"]
    np.testing.assert_allclose(du[0, 1:-1, :], taux / (1000.0 * Hu) * dx * mu[1:-1, :], rtol=1e-14)
    np.testing.assert_allclose(dv[0, :, 1:-1], tauy / (1000.0 * Hv) * dy * mv[:, 1:-1], rtol=1e-14)
    assert np.all(du[0, [0, -1], :] == 0) and np.all(dv[0, :, [0, -1]] == 0)
    assert np.all(du[1:] == 0) and np.all(dv[1:] == 0) and np.all(dh == 0)


def _diffusion_oracle(nx=11, ny=9, mask=None, visc=0.0, diff=0.0):
    f = np.full((nx + 1, ny + 1), 1e-4)
    return SwOracle(nx, ny, 6e3, 9e3, np.array([500., 1000.]), np.array([0.02, 0.01]), f, 120.0, mask=mask,
                    visc_coef=visc, diff_coef=diff)


def test_thickness_diffusion_closed_forms():
    """add_h_diffusion (sw.py:548-565) on h / area = a i^2 + b j^2 (index
    units): 2a / dx^2 + 2b / dy^2 inside; at the first cell of a direction
    the edge padding leaves a / dx^2 (b / dy^2), at the last one
    -a (2n - 3) / dx^2; times kappa area.  Summation by parts with the
    edge padding: sum(q lap q) = -sum(diff_x q^2) / dx^2 - sum(diff_y q^2) / dy^2
    for any q, so zero padding or a dropped boundary term fails; the total
    diffusion tendency is zero (mass conserved)."""
    nx, ny, kap = 11, 9, 250.0
    o = _diffusion_oracle(nx, ny, diff=kap)
    dx, dy, area = o.dx, o.dy, o.area
    a, b = 0.3, -0.7
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    q = a * i ** 2 + b * j ** 2
    got = o.h_diffusion((q * area)[None])[0] / (kap * area)
    lx = np.full(nx, 2 * a / dx ** 2); lx[0] = a / dx ** 2; lx[-1] = -a * (2 * nx - 3) / dx ** 2
    ly = np.full(ny, 2 * b / dy ** 2); ly[0] = b / dy ** 2; ly[-1] = -b * (2 * ny - 3) / dy ** 2
    np.testing.assert_allclose(got, lx[:, None] + ly[None, :], rtol=1e-12, atol=1e-22)
    rng = np.random.default_rng(2)
    q = rng.standard_normal((2, nx, ny))
    lap = o.h_diffusion(q * area) / (kap * area)
    lhs = np.sum(q * lap, axis=(1, 2))
    rhs = -np.sum(np.diff(q, axis=1) ** 2, axis=(1, 2)) / dx ** 2 - np.sum(np.diff(q, axis=2) ** 2, axis=(1, 2)) / dy ** 2
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12)
    assert abs(np.sum(lap)) <= 1e-12 * np.sum(np.abs(lap))


def test_viscosity_closed_forms():
    """add_diffusion (sw.py:511-546) on u / dx = c i^2 + d j^2 and
    v / dy = c i^2 + d j^2 (index units of each grid): on the interior u
    points the x part is 2c / dx^2 everywhere (the boundary u rows are the
    neighbours) and the y part is edge-padded (d / dy^2 at j = 0,
    -d (2 ny - 3) / dy^2 at the last row); on the v points the other way
    round; times nu dx (u) or nu dy (v)."""
    nx, ny, nu = 11, 9, 400.0
    o = _diffusion_oracle(nx, ny, visc=nu)
    dx, dy = o.dx, o.dy
    c, d = 0.4, -0.25
    iu, ju = np.meshgrid(np.arange(nx + 1), np.arange(ny), indexing="ij")
    iv, jv = np.meshgrid(np.arange(nx), np.arange(ny + 1), indexing="ij")
    u = ((c * iu ** 2 + d * ju ** 2) * dx)[None]
    v = ((c * iv ** 2 + d * jv ** 2) * dy)[None]
    du, dv = o.viscosity(u, v)
    ly = np.full(ny, 2 * d / dy ** 2); ly[0] = d / dy ** 2; ly[-1] = -d * (2 * ny - 3) / dy ** 2
    np.testing.assert_allclose(du[0] / (nu * dx), 2 * c / dx ** 2 + ly[None, :] + 0 * du[0], rtol=1e-12)
    lx = np.full(nx, 2 * c / dx ** 2); lx[0] = c / dx ** 2; lx[-1] = -c * (2 * nx - 3) / dx ** 2
    np.testing.assert_allclose(dv[0] / (nu * dy), lx[:, None] + 2 * d / dy ** 2 + 0 * dv[0], rtol=1e-12)
    assert du.shape == (1, nx - 1, ny) and dv.shape == (1, nx, ny - 1)


def test_diffusion_terms_enter_the_tendency_masked():
    """The viscosity is added after the bottom drag on the interior u / v
    points and the whole tendency is masked (sw.py:504-509); the thickness
    diffusion is added to the masked advection and masked itself
    (:743, :564); a coefficient <= 0 adds nothing (:517, :555)."""
    nx, ny = 12, 10
    mask = np.ones((nx, ny)); mask[4:6, 3:5] = 0
    rng = np.random.default_rng(8)
    base = _diffusion_oracle(nx, ny, mask=mask)
    m = base.m
    x = [0.05 * rng.standard_normal((2, nx + 1, ny)) * m["u"], 0.05 * rng.standard_normal((2, nx, ny + 1)) * m["v"],
         0.3 * rng.standard_normal((2, nx, ny)) * m["h"]]
    o = _diffusion_oracle(nx, ny, mask=mask, visc=300.0, diff=200.0)
    r0 = base.rhs(*x)
    r1 = o.rhs(*x)
    vu, vv = o.viscosity(x[0], x[1])
    np.testing.assert_allclose(r1[0] - r0[0], np.pad(vu, ((0, 0), (1, 1), (0, 0))) * m["u"], rtol=1e-9, atol=1e-18)
    np.testing.assert_allclose(r1[1] - r0[1], np.pad(vv, ((0, 0), (0, 0), (1, 1))) * m["v"], rtol=1e-9, atol=1e-18)
    np.testing.assert_allclose(r1[2] - r0[2], o.h_diffusion(x[2]), rtol=1e-9, atol=1e-12)
    assert np.all(o.h_diffusion(x[2])[:, mask == 0] == 0)
    neg = _diffusion_oracle(nx, ny, mask=mask, visc=-5.0, diff=-5.0)
    for a, b in zip(neg.rhs(*x), r0):
        np.testing.assert_array_equal(a, b)
