"""NumPy restatement of the reference's multilayer shallow-water model.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``mapping/models/model_qgsw/`` of datlas-ocean/VarDyn:

* ``sw_masks``         -- masks.py:28-159 (u/v/w masks, vorticity boundary
  classes, the 2/4/6-point stencil masks of the four flux directions)
* ``avg_pool2d``       -- tools.py:9-67 (zero padding, divisor = window area
  unless overridden)
* ``linear2_centered`` reconstruction.py:21-28, ``wenoz4_left`` :82-104,
  ``wenoz6_left`` :135-166, ``smooth_abs`` :15-17 (``WENO_EPS`` = 1e-4, :14)
* ``flux``             -- flux.py:39-90 (velocity-sign biased fluxes with the
  2-, 4- and 6-point reconstructions zero-padded at the ends and blended by the
  stencil masks)
* ``interp_TP`` / ``comp_ke`` / ``div_nofluxbc`` -- finite_diff.py:5-7, 26-30,
  42-52 (edge padding of the face fluxes: the first and last cell of each
  direction get zero divergence)
* ``SwOracle``         -- sw.py:67-937: reference thicknesses (:325-349),
  Coriolis on the u/v grids (:205-215), ``compute_omega`` (:653-671, slip
  boundary classes), ``compute_diagnostic_variables`` (:673-698),
  ``advection_h`` (:465-479), ``advection_momentum`` (:481-510, vortex force,
  Coriolis, kinetic energy + pressure gradient, bottom drag), RK3-SSP
  ``step`` (:752-904, with the external forcing Fu / Fv / Fh added after
  every step, :875-877) with ``set_input_uvh`` / ``get_physical_uvh``
  (:423-449)

Scope (the options of the reference's default configuration,
config_default.py:500-560): per-layer constant depths ``H``, scalar ``dx`` /
``dy``, any binary ocean mask, any ``f`` on the vorticity grid, slip
coefficient, bottom drag, wind stress with a given ``h_wind``
(``add_wind_forcing``, :567-642), Laplacian viscosity of the velocities
(``add_diffusion``, :511-546, its metrics constant: the u/v-grid dx and dy
equal the scalar ones) and of the thickness (``add_h_diffusion``,
:548-565), both with edge-padded (zero-flux) boundaries and only for a
coefficient > 0 as there.  Not restated: sponge, open boundaries and the
barotropic filter (all off by default there; the GPU model rejects them).

The reference quirks are kept: the vorticity flux along y on the u grid
blends the 6-point reconstruction with ``u_sten_wy_gt4`` (so where the
4-point stencil applies both reconstructions are added, sw.py:285-293), the
no-flux divergence comes from edge-padded fluxes (finite_diff.py:42-52), and
``h_tot`` on the u/v grids (sw.py:691-697) is left out because nothing in
this configuration reads it.

Every function works in complex arithmetic, so ``step_tgl`` is the
complex-step derivative of ``step`` (exact to rounding, no step-size
error): the velocity sign tests use the real part, which gives the
derivative of ``max(u, 0)`` as 1 for u > 0 and 0 otherwise.

Pinning (tests/test_oracle_sw.py): masks by a hand-derived known answer for
a rectangular basin and by counting ocean cells in the stencil windows for a
basin with land; reconstructions exact on linear data (both WENO-Z orders,
both directions) and the 6-point one on cell averages of a quadratic; the
rest state stays at rest; Coriolis acceleration of a uniform flow; the
pressure gradient of a thickness ramp in layer 0 or 1 (layer coupling); the
advection of a ramp with zero divergence in the first and last cell;
vorticity of linear shears and at no-slip / partial / free-slip walls; the
mirror symmetries in x and y (with f -> -f); complex-step tangents against
central finite differences; the RK3-SSP amplification polynomial; the
viscosity and thickness diffusion by closed forms on quadratics (which fix
the edge padding and the metrics, dx != dy), summation by parts, mass
conservation and where they enter the tendency.  Parity
unpinned: the 6-point blend at the 4-point faces of the vorticity flux along
y (restated from sw.py:285-293; it only acts next to land).
"""
from __future__ import annotations

import numpy as np

WENO_EPS = 1e-4


# ---------------------------------------------------------------- tools
def avg_pool2d(x, kernel, padding=(0, 0), divisor=None):
    """tools.py:9-67 on a 2-D array: zero padding, stride 1."""
    kh, kw = kernel
    ph, pw = padding
    xp = np.pad(np.asarray(x, dtype=float), ((ph, ph), (pw, pw)))
    H, W = xp.shape[0] - kh + 1, xp.shape[1] - kw + 1
    s = np.zeros((H, W))
    for a in range(kh):
        for b in range(kw):
            s += xp[a:a + H, b:b + W]
    return s / (divisor if divisor is not None else kh * kw)


def sw_masks(mask_h):
    """masks.py:64-159 for a binary h-grid mask (nx, ny)."""
    h = np.asarray(mask_h, dtype=float)
    m = {"h": h}
    u = avg_pool2d(h, (2, 1), (1, 0)) > 3 / 4
    v = avg_pool2d(h, (1, 2), (0, 1)) > 3 / 4
    w = avg_pool2d(h, (2, 2), (1, 1)) > 1 / 8
    one_v = avg_pool2d(v, (2, 1), (1, 0), divisor=1) == 1
    one_u = avg_pool2d(u, (1, 2), (0, 1), divisor=1) == 1
    vb = w & (one_v & ~one_u)
    hb = w & (one_u & ~one_v)
    co = w & (one_u & one_v)
    m.update(u=u, v=v, w=w, w_vertical_bound=vb, w_horizontal_bound=hb, w_cornerout_bound=co,
             w_valid=w & ~vb & ~hb & ~co)
    wf = w.astype(float)
    # h stencil along x on the u grid
    eq2 = (avg_pool2d(h, (2, 1), (1, 0)) > 3 / 4) & (avg_pool2d(h, (4, 1), (2, 0)) < 7 / 8) & u
    gt4 = ~eq2 & u
    eq4 = (avg_pool2d(h, (6, 1), (3, 0)) < 11 / 12) & gt4
    m.update(u_sten_hx_eq2=eq2, u_sten_hx_gt4=gt4, u_sten_hx_eq4=eq4, u_sten_hx_gt6=~eq4 & gt4)
    # w stencil along y on the u grid
    eq2 = (avg_pool2d(wf, (1, 2), (0, 0)) > 3 / 4) & (avg_pool2d(wf, (1, 4), (0, 1)) < 7 / 8) & u
    gt4 = ~eq2 & u
    eq4 = (avg_pool2d(wf, (1, 6), (0, 2)) < 11 / 12) & gt4
    m.update(u_sten_wy_eq2=eq2, u_sten_wy_gt4=gt4, u_sten_wy_eq4=eq4, u_sten_wy_gt6=~eq4 & gt4)
    # h stencil along y on the v grid
    eq2 = (avg_pool2d(h, (1, 2), (0, 1)) > 3 / 4) & (avg_pool2d(h, (1, 4), (0, 2)) < 7 / 8) & v
    gt4 = ~eq2 & v
    eq4 = (avg_pool2d(h, (1, 6), (0, 3)) < 11 / 12) & gt4
    m.update(v_sten_hy_eq2=eq2, v_sten_hy_gt4=gt4, v_sten_hy_eq4=eq4, v_sten_hy_gt6=~eq4 & gt4)
    # w stencil along x on the v grid
    eq2 = (avg_pool2d(wf, (2, 1), (0, 0)) > 3 / 4) & (avg_pool2d(wf, (4, 1), (1, 0)) < 7 / 8) & v
    gt4 = ~eq2 & v
    eq4 = (avg_pool2d(wf, (6, 1), (2, 0)) < 11 / 12) & gt4
    m.update(v_sten_wx_eq2=eq2, v_sten_wx_gt4=gt4, v_sten_wx_eq4=eq4, v_sten_wx_gt6=~eq4 & gt4)
    return {k: np.asarray(a, dtype=float) for k, a in m.items()}


# ---------------------------------------------------------------- reconstructions
def smooth_abs(x):
    return np.sqrt(x ** 2 + WENO_EPS ** 2)


def linear2_centered(qm, qp):
    return 0.5 * (qm + qp)


def wenoz4_left(qmm, qm, qp, qpp):
    qi1 = -1. / 2. * qmm + 3. / 2. * qm
    qi2 = 1. / 2. * (qm + qp)
    beta1 = (qm - qmm) ** 2
    beta2 = (qp - qm) ** 2
    tau = smooth_abs(beta2 - beta1)
    w1 = (1. / 3.) * (1. + tau / (beta1 + WENO_EPS))
    w2 = (2. / 3.) * (1. + tau / (beta2 + WENO_EPS))
    return (w1 * qi1 + w2 * qi2) / (w1 + w2)


def wenoz6_left(qmmm, qmm, qm, qp, qpp, qppp):
    qi1 = 1. / 3. * qmmm - 7. / 6. * qmm + 11. / 6. * qm
    qi2 = -1. / 6. * qmm + 5. / 6. * qm + 1. / 3. * qp
    qi3 = 1. / 3. * qm + 5. / 6. * qp - 1. / 6. * qpp
    k1, k2 = 13. / 12., 0.25
    beta1 = k1 * (qmmm - 2 * qmm + qm) ** 2 + k2 * (qmmm - 4 * qmm + 3 * qm) ** 2
    beta2 = k1 * (qmm - 2 * qm + qp) ** 2 + k2 * (qmm - qp) ** 2
    beta3 = k1 * (qm - 2 * qp + qpp) ** 2 + k2 * (3 * qm - 4 * qp + qpp) ** 2
    tau5 = smooth_abs(beta1 - beta3)
    w1 = 0.1 * (1 + tau5 / (beta1 + WENO_EPS))
    w2 = 0.6 * (1 + tau5 / (beta2 + WENO_EPS))
    w3 = 0.3 * (1 + tau5 / (beta3 + WENO_EPS))
    return (w1 * qi1 + w2 * qi2 + w3 * qi3) / (w1 + w2 + w3)


def _pos(u):
    return np.where(np.real(u) > 0, u, 0 * u)


def flux(q, vel, axis, m2, m4, m6):
    """flux.py:39-90 along ``axis`` (the last two axes are x, y)."""
    q = np.moveaxis(q, axis, -1)
    vel, m2, m4, m6 = (np.moveaxis(np.broadcast_to(a, np.broadcast_shapes(np.shape(a), vel.shape)), axis, -1)
                       for a in (vel, m2, m4, m6))
    n = q.shape[-1]
    up = _pos(vel)
    um = vel - up
    s = [q[..., i:n - 1 + i] for i in range(2)]
    q2 = linear2_centered(*s)
    pad1 = [(0, 0)] * (q.ndim - 1) + [(1, 1)]
    pad2 = [(0, 0)] * (q.ndim - 1) + [(2, 2)]
    s4 = [q[..., i:n - 3 + i] for i in range(4)]
    q4p = np.pad(wenoz4_left(*s4), pad1)
    q4m = np.pad(wenoz4_left(*s4[::-1]), pad1)
    s6 = [q[..., i:n - 5 + i] for i in range(6)]
    q6p = np.pad(wenoz6_left(*s6), pad2)
    q6m = np.pad(wenoz6_left(*s6[::-1]), pad2)
    out = up * (m2 * q2 + m4 * q4p + m6 * q6p) + um * (m2 * q2 + m4 * q4m + m6 * q6m)
    return np.moveaxis(out, -1, axis)


# ---------------------------------------------------------------- finite differences
def interp_TP(f):
    return 0.25 * (f[..., 1:, 1:] + f[..., 1:, :-1] + f[..., :-1, 1:] + f[..., :-1, :-1])


def comp_ke(u, U, v, V):
    usq, vsq = u * U, v * V
    return 0.25 * (usq[..., 1:, :] + usq[..., :-1, :] + vsq[..., 1:] + vsq[..., :-1])


def div_nofluxbc(fx, fy):
    pad_y = [(0, 0)] * (fy.ndim - 1) + [(1, 1)]
    pad_x = [(0, 0)] * (fx.ndim - 2) + [(1, 1), (0, 0)]
    return np.diff(np.pad(fy, pad_y, mode="edge"), axis=-1) + np.diff(np.pad(fx, pad_x, mode="edge"), axis=-2)


def softplus(z):
    """jax.nn.softplus = logaddexp(z, 0), written for complex arguments."""
    big = np.real(z) > 0
    zz = np.where(big, -z, z)
    return np.where(big, z, 0 * z) + np.log1p(np.exp(zz))


def smooth_clamp(x, x_min, sharpness):
    """sw.py:25-31."""
    return x_min + softplus((x - x_min) * sharpness) / sharpness


# ---------------------------------------------------------------- model
class SwOracle:
    """sw.py:67-937 restated for the scope in the module docstring.

    Arrays are (nl, nx+1, ny) for u, (nl, nx, ny+1) for v, (nl, nx, ny) for h
    (x before y, as the reference).  ``f`` lives on the (nx+1, ny+1) vorticity
    grid."""

    def __init__(self, nx, ny, dx, dy, H, g_prime, f, dt, mask=None, slip_coef=0.0,
                 bottom_drag_coef=0.0, h_min=0.1, h_min_sharpness=10.0, taux=0.0, tauy=0.0,
                 h_wind=None, rho_water=1025.0, visc_coef=0.0, diff_coef=0.0):
        self.nx, self.ny, self.nl = nx, ny, len(H)
        self.dx, self.dy = float(dx), float(dy)
        self.area = self.dx * self.dy
        self.H = np.asarray(H, dtype=float)
        self.gp = np.asarray(g_prime, dtype=float)
        self.f = np.asarray(f, dtype=float)
        assert self.f.shape == (nx + 1, ny + 1)
        self.dt = float(dt)
        self.slip = float(slip_coef)
        self.drag = float(bottom_drag_coef)
        self.visc = float(visc_coef or 0.0)
        self.diff = float(diff_coef or 0.0)
        self.h_min, self.sharp = float(h_min), float(h_min_sharpness)
        self.m = sw_masks(np.ones((nx, ny)) if mask is None else mask)
        # reference thicknesses (sw.py:325-349); constant depths give
        # constant h_ref and no reference pressure gradient
        self.h_ref = (self.H * self.area)[:, None, None]
        f_u = 0.5 * (self.f[:, 1:] + self.f[:, :-1])
        f_v = 0.5 * (self.f[1:, :] + self.f[:-1, :])
        self.fstar_u = f_u * self.area
        self.fstar_v = f_v * self.area
        # add_wind_forcing (sw.py:567-642) with h_wind given: H_ref from the
        # scalar or the (nx, ny) field averaged onto the interior u / v points
        self.wind_u = self.wind_v = None
        if np.any(np.asarray(taux) != 0) or np.any(np.asarray(tauy) != 0):
            hw = np.asarray(h_wind, dtype=float)
            if hw.ndim >= 2:
                Hu = np.maximum(0.5 * (hw[:-1, :] + hw[1:, :]), self.h_min)
                Hv = np.maximum(0.5 * (hw[:, :-1] + hw[:, 1:]), self.h_min)
            else:
                Hu = Hv = np.maximum(hw, self.h_min)
            mu, mv = self.m["u"][1:-1, :], self.m["v"][:, 1:-1]
            self.wind_u = np.where(mu > 0.5, taux / (rho_water * Hu) * self.dx, np.zeros_like(mu))
            self.wind_v = np.where(mv > 0.5, tauy / (rho_water * Hv) * self.dy, np.zeros_like(mv))

    # sw.py:653-671
    def omega(self, u, v):
        m = self.m
        u_ = np.pad(u, ((0, 0), (0, 0), (1, 1)))
        v_ = np.pad(v, ((0, 0), (1, 1), (0, 0)))
        dx_v = np.diff(v_, axis=-2)
        dy_u = np.diff(u_, axis=-1)
        curl = dx_v - dy_u
        alpha = 2 * (1 - self.slip)
        return (m["w_valid"] * curl + m["w_cornerout_bound"] * (1 - self.slip) * curl
                + m["w_vertical_bound"] * alpha * dx_v - m["w_horizontal_bound"] * alpha * dy_u)

    @staticmethod
    def lap5(q, dx, dy):
        """Five-point Laplacian of the interior of an array padded by one
        point on both ends of its last two axes (sw.py:533-540, 561-562):
        (q[i+1] - 2 q[i] + q[i-1]) / dx^2 + (q[j+1] - 2 q[j] + q[j-1]) / dy^2."""
        c = q[..., 1:-1, 1:-1]
        return ((q[..., 2:, 1:-1] - 2 * c + q[..., :-2, 1:-1]) / dx ** 2
                + (q[..., 1:-1, 2:] - 2 * c + q[..., 1:-1, :-2]) / dy ** 2)

    def viscosity(self, u, v):
        """add_diffusion (sw.py:511-546): nu lap(u / dx) dx on the interior
        u points and nu lap(v / dy) dy on the interior v points, u edge-padded
        along y and v along x (zero flux through the walls); the other
        neighbours are the boundary rows themselves."""
        up = np.pad(u, ((0, 0), (0, 0), (1, 1)), mode="edge") / self.dx
        vp = np.pad(v, ((0, 0), (1, 1), (0, 0)), mode="edge") / self.dy
        return (self.visc * self.lap5(up, self.dx, self.dy) * self.dx,
                self.visc * self.lap5(vp, self.dx, self.dy) * self.dy)

    def h_diffusion(self, h):
        """add_h_diffusion (sw.py:548-565): kappa lap(h / area) area on
        every cell, h edge-padded on all four sides, masked."""
        hp = np.pad(h / self.area, ((0, 0), (1, 1), (1, 1)), mode="edge")
        return self.diff * self.lap5(hp, self.dx, self.dy) * self.area * self.m["h"]

    def rhs(self, u, v, h):
        """compute_time_derivatives (sw.py:723-750) on scaled variables."""
        m = self.m
        om = self.omega(u, v)
        x = h / self.area
        eta = x + x.sum(axis=0, keepdims=True) - np.cumsum(x, axis=0)
        p = np.cumsum(self.gp[:, None, None] * eta, axis=0)
        U = u / self.dx ** 2
        V = v / self.dy ** 2
        U_m = interp_TP(U)
        V_m = interp_TP(V)
        ke = comp_ke(u, U, v, V) * m["h"]
        # advection_h (sw.py:465-479)
        h_tot = smooth_clamp(self.h_ref + h, self.h_min * self.area, self.sharp)
        fy = flux(h_tot, V[..., 1:-1], -1, m["v_sten_hy_eq2"][:, 1:-1], m["v_sten_hy_eq4"][:, 1:-1],
                  m["v_sten_hy_gt6"][:, 1:-1])
        fx = flux(h_tot, U[..., 1:-1, :], -2, m["u_sten_hx_eq2"][1:-1, :], m["u_sten_hx_eq4"][1:-1, :],
                  m["u_sten_hx_gt6"][1:-1, :])
        dt_h = -div_nofluxbc(fx, fy) * m["h"]
        if self.diff > 0:
            dt_h = dt_h + self.h_diffusion(h)
        # advection_momentum (sw.py:481-510)
        om_Vm = flux(om[..., 1:-1, :], V_m, -1, m["u_sten_wy_eq2"][1:-1, :], m["u_sten_wy_eq4"][1:-1, :],
                     m["u_sten_wy_gt4"][1:-1, :])
        om_Um = flux(om[..., 1:-1], U_m, -2, m["v_sten_wx_eq2"][:, 1:-1], m["v_sten_wx_eq4"][:, 1:-1],
                     m["v_sten_wx_gt6"][:, 1:-1])
        dt_u = om_Vm + self.fstar_u[1:-1, :] * V_m
        dt_v = -(om_Um + self.fstar_v[:, 1:-1] * U_m)
        kep = ke + p
        dt_u = dt_u - (np.diff(kep, axis=-2) + 0.)
        dt_v = dt_v - (np.diff(kep, axis=-1) + 0.)
        # wind forcing of the top layer (sw.py:567-642), then bottom drag (:644-651)
        if self.wind_u is not None:
            dt_u[0] = dt_u[0] + self.wind_u
            dt_v[0] = dt_v[0] + self.wind_v
        dt_u[-1] = dt_u[-1] - self.drag * u[-1, 1:-1, :]
        dt_v[-1] = dt_v[-1] - self.drag * v[-1, :, 1:-1]
        if self.visc > 0:  # after the drag (sw.py:504-506)
            vu, vv = self.viscosity(u, v)
            dt_u = dt_u + vu
            dt_v = dt_v + vv
        dt_u = np.pad(dt_u, ((0, 0), (1, 1), (0, 0))) * m["u"]
        dt_v = np.pad(dt_v, ((0, 0), (0, 0), (1, 1))) * m["v"]
        return dt_u, dt_v, dt_h

    def to_scaled(self, u, v, h):
        """set_input_uvh (sw.py:431-449)."""
        m = self.m
        return (u * m["u"]) * self.dx, (v * m["v"]) * self.dy, (h * m["h"]) * self.area

    def to_physical(self, u, v, h):
        """get_physical_uvh (sw.py:423-429)."""
        return u / self.dx, v / self.dy, h / self.area

    def rk3(self, u, v, h):
        """One RK3-SSP step (sw.py:834-857) on scaled variables."""
        dt = self.dt
        k0 = self.rhs(u, v, h)
        u1, v1, h1 = (a + dt * k for a, k in zip((u, v, h), k0))
        k1 = self.rhs(u1, v1, h1)
        u2, v2, h2 = (a + (dt / 4.0) * (b - 3.0 * c) for a, b, c in zip((u1, v1, h1), k1, k0))
        k2 = self.rhs(u2, v2, h2)
        return tuple(a + (dt / 12.0) * (8.0 * b - c - d) for a, b, c, d in zip((u2, v2, h2), k2, k1, k0))

    def step(self, u0, v0, h0, nstep=1, F=None):
        """step (sw.py:752-904), physical variables in and out.  ``F`` =
        (Fu, Fv, Fh): external forcing per layer in physical units per
        second, scaled like the state and added as x += dt F after every
        RK3 step (sw.py:791-796, 875-877); None: no forcing."""
        s = self.to_scaled(u0, v0, h0)
        Fs = None if F is None else self.to_scaled(*F)
        for _ in range(nstep):
            s = self.rk3(*s)
            if Fs is not None:
                s = tuple(a + self.dt * f for a, f in zip(s, Fs))
        return self.to_physical(*s)

    def step_tgl(self, u0, v0, h0, du0, dv0, dh0, nstep=1, eps=1e-30, F=None, dF=None):
        """Tangent of ``step`` by the complex step (the reference takes
        ``jax.jvp`` of ``step``, sw.py:906-923; with the forcing among the
        differentiated inputs as Model_qgsw.jstep_tgl, mod.py:3222-3253)."""
        z = [np.asarray(a, dtype=complex) + 1j * eps * np.asarray(d) for a, d in
             ((u0, du0), (v0, dv0), (h0, dh0))]
        Fz = None
        if F is not None or dF is not None:
            F = F if F is not None else tuple(np.zeros_like(a) for a in (u0, v0, h0))
            dF = dF if dF is not None else tuple(np.zeros_like(a) for a in F)
            Fz = tuple(np.asarray(a, dtype=complex) + 1j * eps * np.asarray(d) for a, d in zip(F, dF))
        out = self.step(*z, nstep=nstep, F=Fz)
        return tuple(np.imag(o) / eps for o in out)
