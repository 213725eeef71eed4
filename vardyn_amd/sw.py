"""Multilayer rotating shallow-water model on the GPU (class ``SW`` of
mapping/models/model_qgsw/sw.py:67-1035, datlas-ocean/VarDyn).

The same constructor dictionary and the same calls as the reference:
``step(u0, v0, h0, nstep)``, ``step_tgl(u0, v0, h0, du0, dv0, dh0, nstep)``
(the reference's ``jax.jvp`` of ``step``, sw.py:906-923) and
``step_adj(u0, v0, h0, wuT, wvT, whT, nstep)`` (its ``jax.vjp``, sw.py:925-937,
returned in the same nesting: ``((au0, av0, ah0, None, None),)``; the H and
h_wind cotangents are not computed), plus ``adjoint_test_sw`` (:940-1035) and
the physical/scaled conversions.  Fields are the reference's: u (nl, nx+1,
ny), v (nl, nx, ny+1), h (nl, nx, ny), optionally with the ensemble axis of
length 1 in front; f on the (nx+1, ny+1) vorticity grid.

The model runs in libvardyn_hip (``vd_sw_*``, FP64).  Parameters the GPU
model does not implement are rejected rather than ignored: open boundaries,
the barotropic filter, spatially varying depths or metrics, and ensembles of
more than one member.  A ``sponge_coef`` is accepted: the reference's sponge
fields are zero (sw.py:374-377), so its relaxation adds nothing.  Wind stress (with ``h_wind``), the
Laplacian viscosity ``visc_coef`` and the thickness diffusion ``diff_coef``
are built in.  External forcing (Fu, Fv, Fh; the direct momentum mode) is
supported: ``step`` / ``step_tgl`` take it, ``step_adj_forced`` returns its
cotangents.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import SwParams, VardynError
from .qg import current_stream_ptr, ptr


def _scalar(x, what):
    a = np.asarray(x, dtype=np.float64)
    if a.size != 1 and not np.all(a == a.flat[0]):
        raise VardynError(f"sw: {what} must be a scalar (spatially varying {what} is not supported)")
    return float(a.flat[0])


class SW:
    """``SW(param)`` with the reference's parameter dictionary (sw.py:117-323)."""

    _UNSUPPORTED_TRUE = ("barotropic_filter", "flag_obc", "flag_linear")

    def __init__(self, param, device=None):
        p = dict(param)
        self.device = torch.device(device if device is not None else "cuda")
        self.nx, self.ny, self.nl = int(p["nx"]), int(p["ny"]), int(p["nl"])
        self.dx = _scalar(p["dx"], "dx")
        self.dy = _scalar(p["dy"], "dy")
        self.area = self.dx * self.dy
        H = np.asarray(p["H"], dtype=np.float64)
        if H.ndim >= 3 and (H.shape[-1] != 1 or H.shape[-2] != 1):
            raise VardynError("sw: H must be constant per layer (shape (nl, 1, 1))")
        self.H = H.reshape(-1)
        gp = np.asarray(p["g_prime"], dtype=np.float64).reshape(-1)
        if self.H.size != self.nl or gp.size != self.nl:
            raise VardynError(f"sw: H and g_prime need nl = {self.nl} values")
        self.g_prime = gp
        f = np.asarray(p["f"], dtype=np.float64)
        if f.shape != (self.nx + 1, self.ny + 1):
            raise VardynError(f"sw: f must be ({self.nx + 1}, {self.ny + 1}), got {f.shape}")
        self.f = np.ascontiguousarray(f)
        self.dt = float(p["dt"])
        mask = p.get("mask")
        if mask is not None:
            mask = np.asarray(mask)
            if mask.shape != (self.nx, self.ny):
                raise VardynError(f"sw: mask must be ({self.nx}, {self.ny}), got {mask.shape}")
            vals = np.unique(mask)
            if not np.all(np.isin(vals, (0, 1))) or vals.tolist() == [0]:
                raise VardynError("sw: the mask must be binary with at least one ocean cell")
            mask = np.ascontiguousarray(mask, dtype=np.uint8)
        self.mask = mask
        # the sponge (sw.py:825-850) relaxes towards boundary fields with rates
        # sponge_coef / dt times the sponge fields, which the constructor
        # sets to zero (sw.py:374-377) and nothing in the model changes: any
        # sponge_coef leaves the steps unchanged
        self.sponge_coef = float(p.get("sponge_coef", 0.0) or 0.0)
        self.sponge_u = np.zeros((1, 1, self.nx + 1, self.ny))
        self.sponge_v = np.zeros((1, 1, self.nx, self.ny + 1))
        self.sponge_h = np.zeros((1, 1, self.nx, self.ny))
        # add_diffusion / add_h_diffusion (sw.py:511-565): only a positive
        # coefficient diffuses
        self.visc_coef = float(p.get("visc_coef", 0.0) or 0.0)
        self.diff_coef = float(p.get("diff_coef", 0.0) or 0.0)
        for key in self._UNSUPPORTED_TRUE:
            if p.get(key, False):
                raise VardynError(f"sw: {key} is not implemented")
        if int(p.get("n_ens", 1)) != 1:
            raise VardynError("sw: one ensemble member")
        if p.get("forcing_momentum", "direct") != "direct":
            raise VardynError("sw: only direct momentum forcing")
        self.slip_coef = float(p.get("slip_coef", 1.0))
        self.bottom_drag_coef = float(p.get("bottom_drag_coef", 0.0))
        self.h_min = float(p.get("h_min", 0.1))
        self.h_min_sharpness = float(p.get("h_min_sharpness", 10.0))
        self.wind_u, self.wind_v = self._wind(p)
        self._keep = (self.H, self.g_prime, self.f, self.mask, self.wind_u, self.wind_v)
        sp = SwParams(self.nx, self.ny, self.nl, self.dx, self.dy, self.dt,
                      self.H.ctypes.data, self.g_prime.ctypes.data, self.f.ctypes.data,
                      self.mask.ctypes.data if self.mask is not None else None,
                      self.slip_coef, self.bottom_drag_coef, self.h_min, self.h_min_sharpness, self.visc_coef,
                      self.diff_coef,
                      self.wind_u.ctypes.data if self.wind_u is not None else None,
                      self.wind_v.ctypes.data if self.wind_v is not None else None)
        lib = _lib.load()
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(lib.vd_sw_create(C.byref(sp), C.byref(h)), "vd_sw_create")
        self._h = h
        self.u_shape = (self.nl, self.nx + 1, self.ny)
        self.v_shape = (self.nl, self.nx, self.ny + 1)
        self.h_shape = (self.nl, self.nx, self.ny)
        m = _host_masks(self)
        self.mask_u, self.mask_v, self.mask_h = m

    # The reference's sponge fields (sw.py:375-377): zero after construction,
    # which is all this model supports; a caller that sets non-zero fields
    # with a non-zero sponge_coef (Model_qgsw's boundary sponge,
    # mod.py:2694-2760) gets an error instead of a silently ignored sponge.
    def _sponge_get(name):
        return property(lambda self: getattr(self, "_" + name, None),
                        lambda self, val: self._sponge_set(name, val))

    sponge_u = _sponge_get("sponge_u")
    sponge_v = _sponge_get("sponge_v")
    sponge_h = _sponge_get("sponge_h")
    del _sponge_get

    def _sponge_set(self, name, val):
        a = np.asarray(val, dtype=np.float64)
        if self.sponge_coef != 0.0 and np.any(a != 0):
            raise VardynError("sw: non-zero sponge fields are not implemented")
        setattr(self, "_" + name, a)

    def _wind(self, p):
        """add_wind_forcing (sw.py:567-642): the top-layer tendencies
        tau / (rho_water H_ref) dx on the interior u points ((nx-1) x ny) and
        tau / (rho_water H_ref) dy on the interior v points, H_ref from
        ``h_wind`` (a scalar, or an (nx, ny) field averaged onto the u / v
        points) clamped below at ``h_min``; returned on the full u and v grids
        (zero boundary rows), or None without wind.  The kernel multiplies the
        tendency by the u / v masks, so land points get none (the reference's
        ``where``).  With per-layer constant depths the reference's fallback to
        its reference thickness (no ``h_wind``) cannot broadcast, so ``h_wind``
        is required here."""
        nx, ny = self.nx, self.ny
        taux = np.asarray(p.get("taux", 0.0), dtype=np.float64)
        tauy = np.asarray(p.get("tauy", 0.0), dtype=np.float64)
        if not (np.any(taux != 0) or np.any(tauy != 0)):
            return None, None
        if taux.ndim and taux.shape != (nx - 1, ny) or tauy.ndim and tauy.shape != (nx, ny - 1):
            raise VardynError(f"sw: taux must be a scalar or ({nx - 1}, {ny}), tauy a scalar or ({nx}, {ny - 1})")
        if p.get("h_wind") is None:
            raise VardynError("sw: wind stress needs h_wind (the mixed-layer depth, m)")
        rho = float(p.get("rho_water", 1025.0))
        hw = np.asarray(p["h_wind"], dtype=np.float64)
        if hw.ndim >= 2:
            if hw.shape != (nx, ny):
                raise VardynError(f"sw: h_wind must be a scalar or ({nx}, {ny})")
            Hu = np.maximum(0.5 * (hw[:-1, :] + hw[1:, :]), self.h_min)
            Hv = np.maximum(0.5 * (hw[:, :-1] + hw[:, 1:]), self.h_min)
        else:
            Hu = Hv = np.maximum(hw, self.h_min)
        wu = np.zeros((nx + 1, ny))
        wv = np.zeros((nx, ny + 1))
        wu[1:-1, :] = taux / (rho * Hu) * self.dx
        wv[:, 1:-1] = tauy / (rho * Hv) * self.dy
        return np.ascontiguousarray(wu), np.ascontiguousarray(wv)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.load().vd_sw_destroy(h)
            except Exception:
                pass
            self._h = None

    # ------------------------------------------------------------ fields
    def _field(self, a, shape, what):
        t = torch.as_tensor(a, dtype=torch.float64, device=self.device)
        ens = t.dim() == len(shape) + 1
        if ens:
            if t.shape[0] != 1:
                raise VardynError("sw: one ensemble member")
            t = t[0]
        if tuple(t.shape) != shape:
            raise VardynError(f"sw: {what} must be {shape}, got {tuple(t.shape)}")
        return t.contiguous(), ens

    def _state(self, u, v, h, names=("u", "v", "h")):
        (u, e1), (v, e2), (h, e3) = (self._field(u, self.u_shape, names[0]),
                                     self._field(v, self.v_shape, names[1]),
                                     self._field(h, self.h_shape, names[2]))
        return u, v, h, e1 or e2 or e3

    @staticmethod
    def _out(x, ens):
        return x.unsqueeze(0) if ens else x

    # ------------------------------------------------------------ model
    def compute_time_derivatives(self, u, v, h):
        """Time derivatives (dt_u, dt_v, dt_h) at the scaled state (sw.py:723-750)."""
        u, v, h, ens = self._state(u, v, h)
        out = [torch.empty_like(a) for a in (u, v, h)]
        _lib.check(_lib.load().vd_sw_rhs(self._h, ptr(u), ptr(v), ptr(h), *(ptr(o) for o in out),
                                         current_stream_ptr()), "vd_sw_rhs")
        return tuple(self._out(o, ens) for o in out)

    def _forcing(self, Fu, Fv, Fh, names=("Fu", "Fv", "Fh")):
        """Forcing fields (physical units per second, per layer) as device
        tensors; None stays None (a zero field)."""
        shapes = (self.u_shape, self.v_shape, self.h_shape)
        return tuple(None if a is None else self._field(a, shp, nm)[0]
                     for a, shp, nm in zip((Fu, Fv, Fh), shapes, names))

    def step(self, u0, v0, h0, H=None, nstep=1, u_b=None, v_b=None, h_b=None, Fu=None, Fv=None, Fh=None,
             taux=None, tauy=None, h_wind=None):
        """nstep RK3-SSP steps (sw.py:752-904); physical fields in and out.
        The forcing (Fu, Fv, Fh) is added as x += dt F after every step.
        The boundary fields u_b, v_b, h_b only enter through the sponge,
        which this model does not have (sponge_coef = 0), so they change
        nothing."""
        if H is not None:
            raise VardynError("sw: a depth perturbation H is not implemented")
        _no_extras(dict(taux=taux, tauy=tauy, h_wind=h_wind))
        u, v, h, ens = self._state(u0, v0, h0)
        F = self._forcing(Fu, Fv, Fh)
        out = [torch.empty_like(a) for a in (u, v, h)]
        _lib.check(_lib.load().vd_sw_step(self._h, ptr(u), ptr(v), ptr(h), *(ptr(x) for x in F),
                                          *(ptr(o) for o in out), int(nstep), current_stream_ptr()), "vd_sw_step")
        return tuple(self._out(o, ens) for o in out)

    def step_tgl(self, u0, v0, h0, du0, dv0, dh0, H=None, dH=None, nstep=1, Fu=None, Fv=None, Fh=None,
                 dFu=None, dFv=None, dFh=None, **kw):
        """Tangent of ``step`` at (u0, v0, h0; F) applied to (du0, dv0, dh0; dF)
        (the reference's ``jax.jvp`` of ``step``, sw.py:906-923; the forcing
        tangent as in Model_qgsw.jstep_tgl, mod.py:3222-3253)."""
        if H is not None or dH is not None:
            raise VardynError("sw: tangents with respect to H are not implemented")
        _no_extras(kw)
        u, v, h, ens = self._state(u0, v0, h0)
        du, dv, dh, _ = self._state(du0, dv0, dh0, ("du0", "dv0", "dh0"))
        F = self._forcing(Fu, Fv, Fh)
        dF = self._forcing(dFu, dFv, dFh, ("dFu", "dFv", "dFh"))
        out = [torch.empty_like(a) for a in (u, v, h)]
        _lib.check(_lib.load().vd_sw_step_tgl(self._h, ptr(u), ptr(v), ptr(h), *(ptr(x) for x in F), ptr(du),
                                              ptr(dv), ptr(dh), *(ptr(x) for x in dF), *(ptr(o) for o in out),
                                              int(nstep), current_stream_ptr()), "vd_sw_step_tgl")
        return tuple(self._out(o, ens) for o in out)

    def step_adj(self, u0, v0, h0, wuT, wvT, whT, H=None, nstep=1, **kw):
        """Adjoint of ``step`` at (u0, v0, h0) applied to (wuT, wvT, whT), as
        the reference's ``vjp_fn(cotangents)``: ``((au0, av0, ah0, None, None),)``."""
        if H is not None:
            raise VardynError("sw: adjoints with respect to H are not implemented")
        _no_extras(kw)
        au, av, ah, _, _, _ = self.step_adj_forced(u0, v0, h0, wuT, wvT, whT, nstep=nstep, forcing_adjoint=False)
        return ((au, av, ah, None, None),)

    def step_adj_forced(self, u0, v0, h0, wuT, wvT, whT, nstep=1, Fu=None, Fv=None, Fh=None,
                        forcing_adjoint=True):
        """Adjoint of the forced ``step`` (Model_qgsw.jstep_adj, mod.py:3255-3290):
        (au0, av0, ah0, aFu, aFv, aFh), the forcing cotangents None unless
        ``forcing_adjoint``."""
        u, v, h, ens = self._state(u0, v0, h0)
        wu, wv, wh, _ = self._state(wuT, wvT, whT, ("wuT", "wvT", "whT"))
        F = self._forcing(Fu, Fv, Fh)
        out = [torch.empty_like(a) for a in (u, v, h)]
        aF = [torch.empty_like(a) for a in (u, v, h)] if forcing_adjoint else [None, None, None]
        _lib.check(_lib.load().vd_sw_step_adj(self._h, ptr(u), ptr(v), ptr(h), *(ptr(x) for x in F), ptr(wu),
                                              ptr(wv), ptr(wh), *(ptr(o) for o in out), *(ptr(x) for x in aF),
                                              int(nstep), current_stream_ptr()), "vd_sw_step_adj")
        return tuple(self._out(o, ens) for o in out) + tuple(None if a is None else self._out(a, ens) for a in aF)

    # ------------------------------------------------------------ conversions
    def get_physical_uvh(self, u, v, h, numpy=False):
        """sw.py:423-429: scaled -> physical."""
        up, vp, hp = (torch.as_tensor(u) / self.dx, torch.as_tensor(v) / self.dy, torch.as_tensor(h) / self.area)
        return tuple(x.cpu().numpy() for x in (up, vp, hp)) if numpy else (up, vp, hp)

    def set_input_uvh(self, u_phys, v_phys, h_phys):
        """sw.py:431-449: physical -> scaled (masked)."""
        u, v, h, _ = self._state(u_phys, v_phys, h_phys)
        return (u * self.mask_u) * self.dx, (v * self.mask_v) * self.dy, (h * self.mask_h) * self.area

    def adjoint_test_sw(self, nstep=1, seed=42):
        """<M dx, y> / <dx, M* y> with masked random vectors scaled by 1e-4
        (sw.py:940-1035; the reference draws them with jax.random, here a
        seeded NumPy generator)."""
        rng = np.random.default_rng(seed)

        def rand(shape, m):
            return torch.as_tensor(rng.standard_normal(shape) * 1e-4, device=self.device) * m
        ms = (self.mask_u, self.mask_v, self.mask_h)
        shapes = (self.u_shape, self.v_shape, self.h_shape)
        x0 = [rand(s, m) for s, m in zip(shapes, ms)]
        dx0 = [rand(s, m) for s, m in zip(shapes, ms)]
        w = [rand(s, m) for s, m in zip(shapes, ms)]
        dx1 = self.step_tgl(*x0, *dx0, nstep=nstep)
        ad = self.step_adj(*x0, *w, nstep=nstep)[0][:3]
        ps1 = sum(float(torch.sum(a * b)) for a, b in zip(dx1, w))
        ps2 = sum(float(torch.sum(a * b)) for a, b in zip(dx0, ad))
        return ps1 / ps2


def _no_extras(kw):
    extra = {k: v for k, v in kw.items() if v is not None}
    if extra:
        raise VardynError(f"sw: {sorted(extra)} not implemented (boundary fields, forcing, wind, h_wind)")


def _host_masks(model):
    """u, v, h masks on the device (masks.py:64-66), for the conversions."""
    h = np.ones((model.nx, model.ny)) if model.mask is None else model.mask.astype(np.float64)
    hp = np.pad(h, ((1, 1), (0, 0)))
    mu = ((hp[1:] + hp[:-1]) / 2 > 3 / 4).astype(np.float64)
    hq = np.pad(h, ((0, 0), (1, 1)))
    mv = ((hq[:, 1:] + hq[:, :-1]) / 2 > 3 / 4).astype(np.float64)
    dev = model.device
    return (torch.as_tensor(mu, device=dev), torch.as_tensor(mv, device=dev), torch.as_tensor(h, device=dev))
