"""FP64 restatement of the reference 1.5-layer QG model with a hand adjoint.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``mapping/models/model_qg1l/jqgm.py`` of datlas-ocean/VarDyn:

* ``dst1`` / ``inverse_elliptic_dst``  -- jqgm.py:75-91
* ``QgmOracle.__init__``                -- jqgm.py:113-221 (grid, Helmholtz
  eigenvalues, mask rings, beta plane)
* ``h2uv`` jqgm.py:238-262, ``h2pv`` :264-293, ``pv2h`` :295-319,
  ``rhs`` :321-375, ``adv`` :377-396, ``euler/rk2/rk3`` :398-477,
  ``one_step`` :480-503, ``step`` :517-557
* ``step_tgl`` / ``step_adj``: the reference obtains them with ``jax.jvp`` /
  ``jax.vjp`` of ``step_jit`` (jqgm.py:559-569).  Here they are written by
  hand, following JAX's differentiation rules for the primitives used:
  ``jnp.where(c, x, y)`` passes the cotangent to the selected branch, so
  ``up = where(uT < 0, 0, uT)`` and ``um = where(uT > 0, 0, uT)`` BOTH carry
  derivative 1 where ``uT == 0``; ``x.at[idx].set(v)`` zeroes the cotangent
  of the overwritten entries.

Pinning (tests/test_oracle_*.py):

* DST-I: equals ``scipy.fft.dst(type=1, norm='ortho')`` (used directly) and
  is checked against the closed-form sine matrix, orthonormality and
  self-inverse.
* Helmholtz eigenvalues: ``h2pv`` applied to a single DST-I mode returns
  ``lambda_kl`` times the mode; ``pv2h o h2pv`` is the identity on such modes.
* Masks: known-answer test of the ring pattern written out by hand.
* ``h2uv``/``h2pv``/``adv``: polynomial exactness (bilinear / quadratic /
  cubic fields give the analytic derivatives) and point-by-point loops on
  7x9 and 8x10 grids (``tests/test_oracle_pins.py``).
* ``rhs`` optional branches (``tests/test_oracle_pins.py``): the beta plane
  by ``f = f0 + b0 y + g2 y^2`` (centred beta exact, row of v checked), PV
  diffusion by the 5-point Laplacian of a cubic, the bathymetry term by the
  upwind derivative of a cubic ``f_mean * bathymetry`` for both wind signs,
  and all of them superposed.
* Time schemes: amplification polynomials on the linear test equation.
* TLM: Taylor remainder O(lambda^2); adjoint: dot-product identity to 1e-12;
  both against finite differences of the forward.
"""
from __future__ import annotations

import numpy as np
from scipy.fft import dstn


def dst1(x):
    """Orthonormal 2-D DST-I (jqgm.py:75-81: ``dstI2D``)."""
    return dstn(x, type=1, norm="ortho")


def inverse_elliptic_dst(f, operator_dst):
    """``S (S f S / Lambda) S`` (jqgm.py:83-91)."""
    return dst1(dst1(f) / operator_dst)


class QgmOracle:
    """Restatement of ``Qgm`` (jqgm.py:108-569) in FP64 NumPy.

    ``dx``/``dy`` may be 2-D arrays (the centre pixel is used, jqgm.py:126-128)
    or scalars together with ``shape=(ny, nx)``.  ``f`` and ``c`` may be
    scalars or arrays (nanmean; an array ``f`` also defines a beta plane,
    jqgm.py:140-146).
    """

    def __init__(self, dx=None, dy=None, dt=None, SSH=None, c=None, Kdiffus=None,
                 upwind=3, g=9.81, f=1e-4, time_scheme="Euler", mdt=None,
                 bathymetry_PV_term=None, shape=None, dtype=np.float64, **kwargs):
        self.dtype = np.dtype(dtype)
        dt_ = self.dtype.type
        if hasattr(dx, "__len__"):
            ny, nx = np.shape(dx)
            self.dx = dt_(dx[int(ny / 2), int(nx / 2)])
            self.dy = dt_(dy[int(ny / 2), int(nx / 2)])
        else:
            ny, nx = shape
            self.dx = dt_(dx)
            self.dy = dt_(dy)
        self.ny, self.nx = ny, nx
        self.dt = dt
        self.g = g
        if hasattr(f, "__len__"):
            f = np.asarray(f, dtype=np.float64)
            self.f = dt_(np.nanmean(f))
            self.beta = (f[2:, :] - f[:-2, :]) / (2 * self.dy)
        else:
            self.f = dt_(f)
            self.beta = None
        if hasattr(c, "__len__"):
            self.c = dt_(np.nanmean(c))
        else:
            self.c = dt_(c)
        self.mdt = mdt
        self.bathymetry_PV_term = bathymetry_PV_term
        self.upwind = upwind
        if upwind != 3:
            raise NotImplementedError("only the 3rd-order upwind scheme exists in the reference (jqgm.py:377-396)")
        self.time_scheme = time_scheme
        if time_scheme not in ("Euler", "rk2", "rk3"):
            raise ValueError(time_scheme)
        self.Kdiffus = Kdiffus

        # Helmholtz eigenvalues (jqgm.py:165-172)
        x, y = np.meshgrid(np.arange(1, nx - 1, dtype=self.dtype),
                           np.arange(1, ny - 1, dtype=self.dtype))
        laplace_dst = 2 * (np.cos(np.pi / (nx - 1) * x) - 1) / self.dx ** 2 + \
            2 * (np.cos(np.pi / (ny - 1) * y) - 1) / self.dy ** 2
        self.helmoltz_dst = self.g / self.f * laplace_dst - self.g * self.f / self.c ** 2

        self.mask = build_mask(ny, nx, SSH)
        self.ind0 = self.mask == 0
        self.ind1 = self.mask == 1
        self.ind2 = self.mask == 2
        self.ind12 = self.ind1 + self.ind2

    # ------------------------------------------------------------------ forward
    def zeros(self):
        return np.zeros((self.ny, self.nx), dtype=self.dtype)

    def h2uv(self, h):
        """jqgm.py:238-262."""
        g, f = self.g, self.f
        u = self.zeros()
        v = self.zeros()
        with np.errstate(invalid="ignore"):
            u[1:-1, 1:] = -g / f * (h[2:, :-1] + h[2:, 1:] - h[:-2, 1:] - h[:-2, :-1]) / (4 * self.dy)
            v[1:, 1:-1] = g / f * (h[1:, 2:] + h[:-1, 2:] - h[:-1, :-2] - h[1:, :-2]) / (4 * self.dx)
        u = np.where(np.isnan(u), 0, u).astype(self.dtype)
        v = np.where(np.isnan(v), 0, v).astype(self.dtype)
        return u, v

    def h2pv(self, h, hb):
        """jqgm.py:264-293."""
        g, f, c = self.g, self.f, self.c
        q = self.zeros()
        with np.errstate(invalid="ignore"):
            q[1:-1, 1:-1] = g / f * ((h[2:, 1:-1] + h[:-2, 1:-1] - 2 * h[1:-1, 1:-1]) / self.dy ** 2 +
                                     (h[1:-1, 2:] + h[1:-1, :-2] - 2 * h[1:-1, 1:-1]) / self.dx ** 2) - \
                g * f / (c ** 2) * h[1:-1, 1:-1]
        q = np.where(np.isnan(q), 0, q).astype(self.dtype)
        q[self.ind12] = -g * f / (c ** 2) * hb[self.ind12]
        q[self.ind0] = 0
        return q

    def pv2h(self, q, hb, qb):
        """jqgm.py:295-319."""
        qin = q[1:-1, 1:-1] - qb[1:-1, 1:-1]
        h = self.zeros()
        h[1:-1, 1:-1] = inverse_elliptic_dst(qin, self.helmoltz_dst)
        return h + hb

    def _upwind(self, u, v):
        uT = 0.5 * (u[1:-1, 1:-1] + u[1:-1, 2:])
        vT = 0.5 * (v[1:-1, 1:-1] + v[2:, 1:-1])
        up = np.where(uT < 0, 0, uT)
        um = np.where(uT > 0, 0, uT)
        vp = np.where(vT < 0, 0, vT)
        vm = np.where(vT > 0, 0, vT)
        return uT, vT, up, vp, um, vm

    def adv(self, up, vp, um, vm, q0):
        """3rd-order upwind advection (jqgm.py:377-396)."""
        dx, dy = self.dx, self.dy
        out = np.zeros_like(q0, dtype=self.dtype)
        out[2:-2, 2:-2] = (
            - up[1:-1, 1:-1] / (6 * dx) *
            (2 * q0[2:-2, 3:-1] + 3 * q0[2:-2, 2:-2] - 6 * q0[2:-2, 1:-3] + q0[2:-2, :-4])
            + um[1:-1, 1:-1] / (6 * dx) *
            (q0[2:-2, 4:] - 6 * q0[2:-2, 3:-1] + 3 * q0[2:-2, 2:-2] + 2 * q0[2:-2, 1:-3])
            - vp[1:-1, 1:-1] / (6 * dy) *
            (2 * q0[3:-1, 2:-2] + 3 * q0[2:-2, 2:-2] - 6 * q0[1:-3, 2:-2] + q0[:-4, 2:-2])
            + vm[1:-1, 1:-1] / (6 * dy) *
            (q0[4:, 2:-2] - 6 * q0[3:-1, 2:-2] + 3 * q0[2:-2, 2:-2] + 2 * q0[1:-3, 2:-2]))
        return out

    def rhs(self, u, v, q):
        """PV tendency (jqgm.py:321-375) with way=+1."""
        r = self._rhs_raw(u, v, q)
        r = np.where(np.isnan(r), 0, r).astype(self.dtype)
        r[self.ind12] = 0
        r[self.ind0] = 0
        return r

    def _rhs_raw(self, u, v, q):
        """``rhs`` before its NaN filter and mask overwrites."""
        _, _, up, vp, um, vm = self._upwind(u, v)
        r = self.adv(up, vp, um, vm, q)
        if self.bathymetry_PV_term is not None:
            r = r + self.adv(up, vp, um, vm, self.f * self.bathymetry_PV_term)
        if self.beta is not None:
            r[2:-2, 2:-2] = r[2:-2, 2:-2] - self.beta[1:-1, 2:-2] * (v[2:-2, 2:-2] + v[3:-1, 2:-2]) / 2
        if self.Kdiffus is not None:
            K = self.Kdiffus
            r[2:-2, 2:-2] = r[2:-2, 2:-2] + \
                K / (self.dx ** 2) * (q[2:-2, 3:-1] + q[2:-2, 1:-3] - 2 * q[2:-2, 2:-2]) + \
                K / (self.dy ** 2) * (q[3:-1, 2:-2] + q[1:-3, 2:-2] - 2 * q[2:-2, 2:-2])
        return r

    def _k(self, h, q):
        u, v = self.h2uv(h)
        return self.rhs(u, v, q)

    def one_step(self, h0, q0, hb, qb, tape=None):
        """jqgm.py:480-503 together with euler/rk2/rk3 (jqgm.py:398-477).

        ``tape`` (a list) receives the stage states ``(h, q)`` at which ``rhs``
        is evaluated, in order.
        """
        dt = self.dt
        k1 = self._k(h0, q0)
        if tape is not None:
            tape.append((h0, q0))
        if self.time_scheme == "Euler":
            q1 = q0 + dt * k1
        elif self.time_scheme == "rk2":
            q12 = q0 + 0.5 * k1 * dt
            h12 = self.pv2h(q12, hb, qb)
            if tape is not None:
                tape.append((h12, q12))
            k2 = self._k(h12, q12)
            q1 = q0 + dt * k2
        else:
            q12 = q0 + 0.5 * dt * k1
            h12 = self.pv2h(q12, hb, qb)
            if tape is not None:
                tape.append((h12, q12))
            k2 = self._k(h12, q12)
            q13 = q0 - dt * k1 + 2 * dt * k2
            h13 = self.pv2h(q13, hb, qb)
            if tape is not None:
                tape.append((h13, q13))
            k3 = self._k(h13, q13)
            q1 = q0 + (dt / 6) * (k1 + 4 * k2 + k3)
        h1 = self.pv2h(q1, hb, qb)
        return h1, q1

    def step(self, h0, hb, nstep=1, tape=None):
        """jqgm.py:517-557.  ``tape`` receives per-step stage lists."""
        h0 = np.asarray(h0, dtype=self.dtype)
        hb = np.asarray(hb, dtype=self.dtype)
        if self.mdt is not None:
            h0 = h0 + self.mdt
            hb = hb + self.mdt
        q0 = self.h2pv(h0, hb)
        qb = self.h2pv(hb, hb)
        h1 = +h0
        q1 = +q0
        for _ in range(nstep):
            st = [] if tape is not None else None
            h1, q1 = self.one_step(h1, q1, hb, qb, tape=st)
            if tape is not None:
                tape.append(st)
        h1 = h1.copy()
        h1[self.ind0] = np.nan
        if self.mdt is not None:
            h1 = h1 - self.mdt
        return h1

    # ------------------------------------------------------------- linearised
    def _nanmask_uv(self, h):
        with np.errstate(invalid="ignore"):
            u = np.zeros((self.ny, self.nx))
            v = np.zeros((self.ny, self.nx))
            u[1:-1, 1:] = (h[2:, :-1] + h[2:, 1:] - h[:-2, 1:] - h[:-2, :-1])
            v[1:, 1:-1] = (h[1:, 2:] + h[:-1, 2:] - h[:-1, :-2] - h[1:, :-2])
        return np.isnan(u), np.isnan(v)

    def _nanmask_q(self, h):
        with np.errstate(invalid="ignore"):
            q = np.zeros((self.ny, self.nx))
            q[1:-1, 1:-1] = (h[2:, 1:-1] + h[:-2, 1:-1] + h[1:-1, 2:] + h[1:-1, :-2] + h[1:-1, 1:-1])
        return np.isnan(q)

    def h2uv_tl(self, dh, h):
        du, dv = self.zeros(), self.zeros()
        g, f = self.g, self.f
        du[1:-1, 1:] = -g / f * (dh[2:, :-1] + dh[2:, 1:] - dh[:-2, 1:] - dh[:-2, :-1]) / (4 * self.dy)
        dv[1:, 1:-1] = g / f * (dh[1:, 2:] + dh[:-1, 2:] - dh[:-1, :-2] - dh[1:, :-2]) / (4 * self.dx)
        mu, mv = self._nanmask_uv(h)
        du[mu] = 0
        dv[mv] = 0
        return du, dv

    def h2uv_adj(self, adu, adv, h):
        g, f = self.g, self.f
        mu, mv = self._nanmask_uv(h)
        adu = np.where(mu, 0, adu)
        adv = np.where(mv, 0, adv)
        adh = self.zeros()
        a = -g / f * adu[1:-1, 1:] / (4 * self.dy)
        adh[2:, :-1] += a
        adh[2:, 1:] += a
        adh[:-2, 1:] -= a
        adh[:-2, :-1] -= a
        b = g / f * adv[1:, 1:-1] / (4 * self.dx)
        adh[1:, 2:] += b
        adh[:-1, 2:] += b
        adh[:-1, :-2] -= b
        adh[1:, :-2] -= b
        return adh

    def h2pv_tl(self, dh, h):
        g, f, c = self.g, self.f, self.c
        dq = self.zeros()
        dq[1:-1, 1:-1] = g / f * ((dh[2:, 1:-1] + dh[:-2, 1:-1] - 2 * dh[1:-1, 1:-1]) / self.dy ** 2 +
                                  (dh[1:-1, 2:] + dh[1:-1, :-2] - 2 * dh[1:-1, 1:-1]) / self.dx ** 2) - \
            g * f / (c ** 2) * dh[1:-1, 1:-1]
        dq[self._nanmask_q(h)] = 0
        dq[self.ind12] = 0
        dq[self.ind0] = 0
        return dq

    def h2pv_adj(self, adq, h):
        g, f, c = self.g, self.f, self.c
        adq = adq.copy()
        adq[self.ind0] = 0
        adq[self.ind12] = 0
        adq[self._nanmask_q(h)] = 0
        a = adq[1:-1, 1:-1]
        adh = self.zeros()
        cy = g / f / self.dy ** 2
        cx = g / f / self.dx ** 2
        adh[2:, 1:-1] += cy * a
        adh[:-2, 1:-1] += cy * a
        adh[1:-1, 2:] += cx * a
        adh[1:-1, :-2] += cx * a
        adh[1:-1, 1:-1] += (-2 * cy - 2 * cx - g * f / c ** 2) * a
        return adh

    def pv2h_lin(self, dq):
        """Linear part of ``pv2h``; symmetric, so it is also its adjoint."""
        dh = self.zeros()
        dh[1:-1, 1:-1] = inverse_elliptic_dst(dq[1:-1, 1:-1], self.helmoltz_dst)
        return dh

    def rhs_tl(self, du, dv, dq, u, v, q):
        uT, vT, up, vp, um, vm = self._upwind(u, v)
        duT = 0.5 * (du[1:-1, 1:-1] + du[1:-1, 2:])
        dvT = 0.5 * (dv[1:-1, 1:-1] + dv[2:, 1:-1])
        dup = np.where(uT < 0, 0, duT)
        dum = np.where(uT > 0, 0, duT)
        dvp = np.where(vT < 0, 0, dvT)
        dvm = np.where(vT > 0, 0, dvT)
        dr = self.adv(dup, dvp, dum, dvm, q) + self.adv(up, vp, um, vm, dq)
        if self.bathymetry_PV_term is not None:
            dr = dr + self.adv(dup, dvp, dum, dvm, self.f * self.bathymetry_PV_term)
        if self.beta is not None:
            dr[2:-2, 2:-2] = dr[2:-2, 2:-2] - self.beta[1:-1, 2:-2] * (dv[2:-2, 2:-2] + dv[3:-1, 2:-2]) / 2
        if self.Kdiffus is not None:
            K = self.Kdiffus
            dr[2:-2, 2:-2] = dr[2:-2, 2:-2] + \
                K / (self.dx ** 2) * (dq[2:-2, 3:-1] + dq[2:-2, 1:-3] - 2 * dq[2:-2, 2:-2]) + \
                K / (self.dy ** 2) * (dq[3:-1, 2:-2] + dq[1:-3, 2:-2] - 2 * dq[2:-2, 2:-2])
        with np.errstate(invalid="ignore"):
            dr[np.isnan(self._rhs_raw(u, v, q))] = 0
        dr[self.ind12] = 0
        dr[self.ind0] = 0
        return dr

    def adv_adj(self, ad, up, vp, um, vm, q0, want_q=True):
        """Transpose of ``adv`` w.r.t. (up, vp, um, vm) and q0."""
        dx, dy = self.dx, self.dy
        a = ad[2:-2, 2:-2]
        Dxp = (2 * q0[2:-2, 3:-1] + 3 * q0[2:-2, 2:-2] - 6 * q0[2:-2, 1:-3] + q0[2:-2, :-4]) / (6 * dx)
        Dxm = (q0[2:-2, 4:] - 6 * q0[2:-2, 3:-1] + 3 * q0[2:-2, 2:-2] + 2 * q0[2:-2, 1:-3]) / (6 * dx)
        Dyp = (2 * q0[3:-1, 2:-2] + 3 * q0[2:-2, 2:-2] - 6 * q0[1:-3, 2:-2] + q0[:-4, 2:-2]) / (6 * dy)
        Dym = (q0[4:, 2:-2] - 6 * q0[3:-1, 2:-2] + 3 * q0[2:-2, 2:-2] + 2 * q0[1:-3, 2:-2]) / (6 * dy)
        shp = up.shape
        adup = np.zeros(shp)
        adum = np.zeros(shp)
        advp = np.zeros(shp)
        advm = np.zeros(shp)
        adup[1:-1, 1:-1] = -a * Dxp
        adum[1:-1, 1:-1] = a * Dxm
        advp[1:-1, 1:-1] = -a * Dyp
        advm[1:-1, 1:-1] = a * Dym
        adq = None
        if want_q:
            adq = np.zeros_like(q0)
            c1 = -up[1:-1, 1:-1] / (6 * dx) * a
            adq[2:-2, 3:-1] += 2 * c1
            adq[2:-2, 2:-2] += 3 * c1
            adq[2:-2, 1:-3] += -6 * c1
            adq[2:-2, :-4] += c1
            c2 = um[1:-1, 1:-1] / (6 * dx) * a
            adq[2:-2, 4:] += c2
            adq[2:-2, 3:-1] += -6 * c2
            adq[2:-2, 2:-2] += 3 * c2
            adq[2:-2, 1:-3] += 2 * c2
            c3 = -vp[1:-1, 1:-1] / (6 * dy) * a
            adq[3:-1, 2:-2] += 2 * c3
            adq[2:-2, 2:-2] += 3 * c3
            adq[1:-3, 2:-2] += -6 * c3
            adq[:-4, 2:-2] += c3
            c4 = vm[1:-1, 1:-1] / (6 * dy) * a
            adq[4:, 2:-2] += c4
            adq[3:-1, 2:-2] += -6 * c4
            adq[2:-2, 2:-2] += 3 * c4
            adq[1:-3, 2:-2] += 2 * c4
        return adup, advp, adum, advm, adq

    def rhs_adj(self, ar, u, v, q):
        """Transpose of ``rhs_tl``: returns (adu, adv, adq)."""
        uT, vT, up, vp, um, vm = self._upwind(u, v)
        ar = ar.copy()
        ar[self.ind0] = 0
        ar[self.ind12] = 0
        with np.errstate(invalid="ignore"):
            ar[np.isnan(self._rhs_raw(u, v, q))] = 0
        adu = self.zeros()
        adv = self.zeros()
        adup, advp, adum, advm, adq = self.adv_adj(ar, up, vp, um, vm, q)
        if self.bathymetry_PV_term is not None:
            b_up, b_vp, b_um, b_vm, _ = self.adv_adj(ar, up, vp, um, vm, self.f * self.bathymetry_PV_term,
                                                     want_q=False)
            adup += b_up
            advp += b_vp
            adum += b_um
            advm += b_vm
        if self.beta is not None:
            a = ar[2:-2, 2:-2]
            bb = -self.beta[1:-1, 2:-2] * a / 2
            adv[2:-2, 2:-2] += bb
            adv[3:-1, 2:-2] += bb
        if self.Kdiffus is not None:
            K = self.Kdiffus
            a = ar[2:-2, 2:-2]
            cx = K / (self.dx ** 2) * a
            cy = K / (self.dy ** 2) * a
            adq[2:-2, 3:-1] += cx
            adq[2:-2, 1:-3] += cx
            adq[2:-2, 2:-2] += -2 * cx - 2 * cy
            adq[3:-1, 2:-2] += cy
            adq[1:-3, 2:-2] += cy
        aduT = np.where(uT < 0, 0, adup) + np.where(uT > 0, 0, adum)
        advT = np.where(vT < 0, 0, advp) + np.where(vT > 0, 0, advm)
        adu[1:-1, 1:-1] += 0.5 * aduT
        adu[1:-1, 2:] += 0.5 * aduT
        adv[1:-1, 1:-1] += 0.5 * advT
        adv[2:, 1:-1] += 0.5 * advT
        return adu, adv, adq

    def _k_tl(self, dh, dq, h, q):
        u, v = self.h2uv(h)
        du, dv = self.h2uv_tl(dh, h)
        return self.rhs_tl(du, dv, dq, u, v, q)

    def _k_adj(self, ak, h, q):
        u, v = self.h2uv(h)
        adu, adv, adq = self.rhs_adj(ak, u, v, q)
        return self.h2uv_adj(adu, adv, h), adq

    def step_tgl(self, dh0, h0, hb, nstep=1):
        """Tangent linear of ``step`` (the reference's ``jvp``, jqgm.py:559-563)."""
        h0 = np.asarray(h0, dtype=self.dtype)
        hb = np.asarray(hb, dtype=self.dtype)
        tape = []
        h1 = self.step(h0, hb, nstep, tape=tape)
        if self.mdt is not None:
            h0 = h0 + self.mdt
        dt = self.dt
        dh = np.asarray(dh0, dtype=self.dtype).copy()
        dq = self.h2pv_tl(dh, h0)
        for st in tape:
            if self.time_scheme == "Euler":
                (ha, qa), = st
                dk1 = self._k_tl(dh, dq, ha, qa)
                dq1 = dq + dt * dk1
            elif self.time_scheme == "rk2":
                (ha, qa), (hb_, qb_) = st
                dk1 = self._k_tl(dh, dq, ha, qa)
                dq12 = dq + 0.5 * dt * dk1
                dh12 = self.pv2h_lin(dq12)
                dk2 = self._k_tl(dh12, dq12, hb_, qb_)
                dq1 = dq + dt * dk2
            else:
                (ha, qa), (hb_, qb_), (hc, qc) = st
                dk1 = self._k_tl(dh, dq, ha, qa)
                dq12 = dq + 0.5 * dt * dk1
                dh12 = self.pv2h_lin(dq12)
                dk2 = self._k_tl(dh12, dq12, hb_, qb_)
                dq13 = dq - dt * dk1 + 2 * dt * dk2
                dh13 = self.pv2h_lin(dq13)
                dk3 = self._k_tl(dh13, dq13, hc, qc)
                dq1 = dq + (dt / 6) * (dk1 + 4 * dk2 + dk3)
            dh = self.pv2h_lin(dq1)
            dq = dq1
        dh = dh.copy()
        dh[self.ind0] = 0
        return dh, h1

    def step_adj(self, adh1, h0, hb, nstep=1):
        """Adjoint of ``step`` (the reference's ``vjp``, jqgm.py:565-569)."""
        h0 = np.asarray(h0, dtype=self.dtype)
        hb = np.asarray(hb, dtype=self.dtype)
        tape = []
        self.step(h0, hb, nstep, tape=tape)
        if self.mdt is not None:
            h0 = h0 + self.mdt
        dt = self.dt
        adh = np.asarray(adh1, dtype=self.dtype).copy()
        adh[self.ind0] = 0
        adq = self.zeros()
        for n in reversed(range(nstep)):
            st = tape[n]
            # h_{n+1} = pv2h(q_{n+1})
            adq = adq + self.pv2h_lin(adh)
            if self.time_scheme == "Euler":
                (ha, qa), = st
                ak1 = dt * adq
                adh, adq_k = self._k_adj(ak1, ha, qa)
                adq = adq + adq_k
            elif self.time_scheme == "rk2":
                (ha, qa), (hm, qm) = st
                ak2 = dt * adq
                adh12, adq12 = self._k_adj(ak2, hm, qm)
                adq12 = adq12 + self.pv2h_lin(adh12)
                adq0 = adq + adq12
                ak1 = 0.5 * dt * adq12
                adh, adq_k = self._k_adj(ak1, ha, qa)
                adq = adq0 + adq_k
            else:
                (ha, qa), (hm, qm), (hc, qc) = st
                ak1 = (dt / 6) * adq
                ak2 = (4 * dt / 6) * adq
                ak3 = (dt / 6) * adq
                adq0 = adq.copy()
                adh13, adq13 = self._k_adj(ak3, hc, qc)
                adq13 = adq13 + self.pv2h_lin(adh13)
                adq0 = adq0 + adq13
                ak1 = ak1 - dt * adq13
                ak2 = ak2 + 2 * dt * adq13
                adh12, adq12 = self._k_adj(ak2, hm, qm)
                adq12 = adq12 + self.pv2h_lin(adh12)
                adq0 = adq0 + adq12
                ak1 = ak1 + 0.5 * dt * adq12
                adh, adq_k = self._k_adj(ak1, ha, qa)
                adq = adq0 + adq_k
        return adh + self.h2pv_adj(adq, h0)


def build_mask(ny, nx, SSH=None):
    """Mask rings of jqgm.py:179-217 (3 interior, 1 outer ring / coast,
    2 next to the border / coast, 0 land)."""
    mask = 3 * np.ones((ny, nx), dtype=np.int64)
    mask[0, :] = 1
    mask[:, 0] = 1
    mask[-1, :] = 1
    mask[:, -1] = 1
    mask[1, 1:-1] = 2
    mask[1:-1, 1] = 2
    mask[-2, 1:-1] = 2
    mask[-3, 1:-1] = 2
    mask[1:-1, -2] = 2
    mask[1:-1, -3] = 2
    if SSH is not None:
        isnan = np.isnan(SSH)
        mask[isnan] = 0
        for i, j in np.argwhere(isnan):
            for p1 in range(-2, 3):
                for p2 in range(-2, 3):
                    it, jt = i + p1, j + p2
                    if 0 <= it <= ny - 1 and 0 <= jt <= nx - 1:
                        if mask[it, jt] >= 2 and (p1 in (-1, 0, 1) and p2 in (-1, 0, 1)):
                            mask[it, jt] = 1
                        elif mask[it, jt] == 3:
                            mask[it, jt] = 2
    return mask
