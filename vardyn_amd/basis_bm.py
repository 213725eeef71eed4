"""The balanced-motion wavelet basis (``Basis_bm``, mapping/src/basis.py:472-936,
datlas-ocean/VarDyn) with its synthesis and transpose on the GPU.

The reduced control of the reference's flux and BM configurations: SSH (or an
SSH forcing) at time t is a sum of space-time wavelets.  For every
pseudo-frequency f (log-spaced between 1/lmin and 1/lmax), direction theta and
space centre P (latitude rows DX_f/facns km apart, longitudes spaced by the
same distance on each row, centres over a fully masked box dropped) there is a
cosine and a sine wave sqrt(2) w(dx/DX_f) w(dy/DX_f) cos|sin(k (dx cos theta +
dy sin theta)) on the centre's box |dx|, |dy| <= DX_f, w(x) = cos^2(pi x / 2);
each repeats at time centres tdec_f/facnlt apart, weighted at time t by
window((t - t_c)/tdec_f) (normalised by the window's integral with
``norm_time``), window = w, or its derivative shape with ``flux``
(basis.py:564-820, 4138-4144).  The coefficient vector is laid out as in the
reference: frequency, time centre, space centre, direction, cos/sin, so
controls, prior variances Q and gradients are interchangeable with it.

Here ``set_basis`` builds the tables (centres, their point boxes on the
rectilinear grid, the centres covering each grid point) and ``operg`` /
``operg_transpose`` run ``vd_bm_operg`` / ``vd_bm_operg_adj``: a matrix-free,
deterministic gather in each direction, for one time or a batch of times
(``operg_batch`` / ``operg_transpose_batch``).  With ``compute_velocities``
the fields also give their geostrophic velocities (basis.py:822-860) through
``vd_bm_ssh2uv`` / ``vd_bm_ssh2uv_adj``.  Scope: no depth attenuation and no
background file (both read netCDF files), no multi-mode accumulation.
"""
from __future__ import annotations

import ctypes as C
import datetime
from types import SimpleNamespace

import numpy as np
import torch

from . import _lib
from ._lib import VardynError
from .qg import current_stream_ptr, ptr

KM2DEG = 1.0 / 110


def _window(x):
    return np.cos(x * 0.5 * np.pi) ** 2


def _window_flux(x):
    return -np.pi * np.sin(x * 0.5 * np.pi) * np.cos(x * 0.5 * np.pi)


class BmParams(C.Structure):
    _fields_ = [("ny", C.c_int), ("nx", C.c_int), ("nf", C.c_int), ("ntheta", C.c_int),
                ("nc", C.c_int)] + [(n, C.c_void_p) for n in (
                    "lon1", "lat1", "mask", "dxf", "kx", "ky", "cf", "cP", "clon", "clat", "ccos",
                    "cbox", "fbase", "fnp", "fnt", "pptr", "plist")]


def _host(a, dtype):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


class Basis_bm:
    """``Basis_bm`` (basis.py:472-936) on the GPU.

    ``State`` needs 2-D ``lon``, ``lat`` (degrees) of a rectilinear grid and
    an optional boolean ``mask`` (True = masked); the keyword arguments are
    the reference's ``config.BASIS`` fields (defaults of BASIS_BM,
    mapping/src/config_default.py:804-860) and ``delta_time_ref`` the days
    from 1950-01-01 to the experiment's start (basis.py:528)."""

    controls_initial_state = False
    time_dependent = True

    def __init__(self, State, flux=False, facns=1.0, facnlt=2.0, npsp=3.5, facpsp=1.5,
                 lmin=80.0, lmax=970.0, lmeso=300.0, tmeso=20.0, sloptdec=-1.28, factdec=0.5,
                 tdecmin=2.5, tdecmax=40.0, facQ=1.0, Qmax=1e-3, slopQ=-5.0, norm_time=True,
                 name_mod_var="ssh", delta_time_ref=0.0, compute_velocities=False, name_mod_u="u",
                 name_mod_v="v", device=None, **kwargs):
        if not torch.cuda.is_available():
            raise VardynError("Basis_bm needs a GPU (libvardyn_hip)")
        lon, lat = np.asarray(State.lon, dtype=np.float64), np.asarray(State.lat, dtype=np.float64)
        if lon.ndim != 2 or lon.shape != lat.shape:
            raise VardynError("Basis_bm: lon and lat must be 2-D arrays of the grid's shape")
        if not (np.all(lon == lon[:1, :]) and np.all(lat == lat[:, :1])):
            raise VardynError("Basis_bm: the grid must be rectilinear (lon along x, lat along y)")
        self.ny, self.nx = lon.shape
        self.shape_phys = (self.ny, self.nx)
        self.lon1, self.lat1 = lon[0].copy(), lat[:, 0].copy()
        self.lon_min, self.lon_max = lon.min(), lon.max()
        self.lat_min, self.lat_max = lat.min(), lat.max()
        m = getattr(State, "mask", None)
        m = None if m is None else np.asarray(m, dtype=bool)
        self.mask = m if (m is not None and np.any(m)) else None
        self.flux, self.facns, self.facnlt, self.npsp, self.facpsp = flux, facns, facnlt, npsp, facpsp
        self.lmin, self.lmax, self.lmeso, self.tmeso = lmin, lmax, lmeso, tmeso
        self.sloptdec, self.factdec, self.tdecmin, self.tdecmax = sloptdec, factdec, tdecmin, tdecmax
        self.Qmax, self.slopQ, self.facQ, self.norm_time = Qmax, slopQ, facQ, norm_time
        self.name_mod_var = name_mod_var
        self.compute_velocities = bool(compute_velocities)
        self.name_mod_u, self.name_mod_v = name_mod_u, name_mod_v
        self._uv = None
        if self.compute_velocities:
            # basis.py:514-526: f, DX, DY padded by one edge row and column,
            # averaged onto the u and v grids; g = 9.81
            try:
                f, DX, DY = (np.asarray(getattr(State, n), dtype=np.float64) * np.ones(lon.shape)
                             for n in ("f", "DX", "DY"))
            except AttributeError as e:
                raise VardynError("Basis_bm with compute_velocities needs State.f, State.DX and State.DY") from e
            pad = ((1, 0), (1, 0))
            _f = np.pad(f, pad_width=pad, mode="edge")
            f_on_v = 0.5 * (_f[:, 1:] + _f[:, :-1])
            f_on_u = 0.5 * (_f[1:, :] + _f[:-1, :])
            dx = np.pad(DX, pad_width=pad, mode="edge")
            dy = np.pad(DY, pad_width=pad, mode="edge")
            self.g = 9.81
            self._uv = tuple(np.ascontiguousarray(a) for a in (
                -self.g / f_on_u, 0.5 * (dy[1:, :] + dy[:-1, :]), self.g / f_on_v,
                0.5 * (dx[:, 1:] + dx[:, :-1])))
        self.delta_time_ref = float(delta_time_ref)
        self.window = _window_flux if flux else _window
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None \
            else torch.device(device)
        self._h = None

    @classmethod
    def from_config(cls, config, State, **kw):
        """The reference's ``Basis_bm(config, State)`` (basis.py:474-562)."""
        B = config.BASIS
        names = ("flux", "facns", "facnlt", "npsp", "facpsp", "lmin", "lmax", "lmeso", "tmeso",
                 "sloptdec", "factdec", "tdecmin", "tdecmax", "facQ", "Qmax", "slopQ",
                 "norm_time", "name_mod_var", "compute_velocities", "name_mod_u", "name_mod_v")
        args = {n: getattr(B, n) for n in names if hasattr(B, n)}
        ref = (config.EXP.init_date - datetime.datetime(1950, 1, 1, 0)).total_seconds() / 24 / 3600
        return cls(State, delta_time_ref=ref, **args, **kw)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            try:
                _lib.load().vd_bm_destroy(h)
            except Exception:
                pass
            self._h = None

    # ------------------------------------------------------------------ set-up
    def _rows_in(self, lat_c, half):
        """Grid rows i with |(lat_i - lat_c) / km2deg| <= half (contiguous)."""
        ok = np.abs((self.lat1 - lat_c) / KM2DEG) <= half
        return np.nonzero(ok)[0]

    def _cols_in(self, lon_c, coslat, half):
        ok = np.abs((self.lon1 - lon_c) / KM2DEG * coslat) <= half
        return np.nonzero(ok)[0]

    def set_basis(self, time, return_q=False, **kwargs):
        """Centres, time structure and prior variances (basis.py:564-724)."""
        time = np.asarray(time, dtype=np.float64)
        lon_min, lon_max, lat_min, lat_max = self.lon_min, self.lon_max, self.lat_min, self.lat_max
        if lon_max < lon_min:
            lon_max = lon_max + 360.0
        step = np.log(1 + self.facpsp / self.npsp)
        ff = np.exp(np.arange(np.log(1.0 / self.lmin), np.log(1.0 / self.lmax) - step, -step)[::-1])
        dff = ff[1:] - ff[:-1]
        theta = np.linspace(0, np.pi, int(np.pi * ff[0] / dff[0] * self.facpsp))[:-1]
        nf, ntheta = len(ff), len(theta)
        span = time.max() - time.min()
        DX = 1.0 / ff * self.npsp * 0.5
        DXG = DX / self.facns
        self.centres = []  # per frequency: (lon, lat) arrays
        self.tdec, self.enst, self.norm_fact = [], [], []
        for f in range(nf):
            dlat = DXG[f] * KM2DEG
            lat0 = lat_min - lat_min % dlat - DX[f] * KM2DEG
            rows_lat = np.arange(lat0, lat_max + DX[f] * KM2DEG * 1.5, dlat)
            lons, lats = [], []
            for la in rows_lat:
                cl = np.cos(la * np.pi / 180.0)
                dlon = DXG[f] / cl * KM2DEG
                lon0 = lon_min - lon_min % dlon - DX[f] / cl * KM2DEG
                row = np.arange(lon0, lon_max + DX[f] / cl * KM2DEG * 1.5, dlon)
                if self.mask is not None:
                    # drop centres whose (1/f) box holds only masked points (basis.py:640-655)
                    ri = self._rows_in(la, 1.0 / ff[f])
                    keep = np.ones(row.size, dtype=bool)
                    for n, lo in enumerate(row):
                        cj = self._cols_in(lo, cl, 1.0 / ff[f])
                        box = self.mask[np.ix_(ri, cj)]
                        keep[n] = not (box.size == 0 or np.all(box))
                    row = row[keep]
                lons.append(row)
                lats.append(np.full(row.size, la))
            self.centres.append((np.concatenate(lons) if lons else np.zeros(0),
                                 np.concatenate(lats) if lats else np.zeros(0)))
            td = self.tmeso * self.lmeso ** self.sloptdec * ff[f] ** self.sloptdec * self.factdec
            td = min(max(td, self.tdecmin), self.tdecmax)
            self.tdec.append(td)
            t0 = -self.delta_time_ref % td
            self.enst.append(np.arange(t0 - td / self.facnlt, span + td / self.facnlt,
                                       td / self.facnlt))
            tt = np.linspace(-td, td)
            cum = np.concatenate([[0.0], np.cumsum(self.window(tt[:-1] / td) * (tt[1:] - tt[:-1]))])
            self.norm_fact.append(cum.max())
        NP = np.array([c[0].size for c in self.centres], dtype=np.int64)
        nt = np.array([e.size for e in self.enst], dtype=np.int64)
        nwf = np.where(NP > 0, 2 * nt * ntheta * NP, 0)
        self.iff_wavebounds = [0] + list(np.cumsum(nwf))
        Q = []
        for f in range(nf):
            if NP[f] > 0:
                q = self.Qmax / self.facns if 1.0 / ff[f] > self.lmeso else \
                    self.Qmax / self.facns * self.lmeso ** self.slopQ * ff[f] ** self.slopQ
                Q.append(np.full(int(nwf[f]), q))
        self.Q = np.concatenate(Q) if Q else np.zeros(0)
        self.nbasis = int(self.Q.size)
        self.ff, self.theta, self.nf, self.ntheta, self.k = ff, theta, nf, ntheta, 2 * np.pi * ff
        self.DX, self.NP, self.nt = DX, NP, nt
        self._build_tables()
        self.time = time
        if return_q:
            return np.zeros_like(self.Q), self.Q

    def _build_tables(self):
        cf, cP, clon, clat, box, pts = [], [], [], [], [], []
        for f in range(self.nf):
            lons, lats = self.centres[f]
            for P in range(lons.size):
                lo, la = lons[P], lats[P]
                cl = np.cos(la * np.pi / 180.0)
                ri = self._rows_in(la, self.DX[f])
                cj = self._cols_in(lo, cl, self.DX[f])
                c = len(cf)
                cf.append(f)
                cP.append(P)
                clon.append(lo)
                clat.append(la)
                if ri.size and cj.size:
                    if ri[-1] - ri[0] + 1 != ri.size or cj[-1] - cj[0] + 1 != cj.size:
                        raise VardynError("Basis_bm: centre box is not contiguous on this grid")
                    box.append((ri[0], ri[-1], cj[0], cj[-1]))
                    p = (ri[:, None] * self.nx + cj[None, :]).ravel()
                    if self.mask is not None:
                        p = p[~self.mask.ravel()[p]]
                    pts.append(np.stack([p, np.full(p.size, c)], axis=1))
                else:
                    box.append((1, 0, 1, 0))
        self.nc = len(cf)
        pairs = np.concatenate(pts) if pts else np.zeros((0, 2), dtype=np.int64)
        order = np.argsort(pairs[:, 0], kind="stable")  # centres ascending within a point
        pairs = pairs[order]
        counts = np.bincount(pairs[:, 0], minlength=self.ny * self.nx)
        self.nnz_space = int(pairs.shape[0]) * 2 * self.ntheta  # entries of the reference's Gx
        self._tab = {
            "lon1": _host(self.lon1, np.float64), "lat1": _host(self.lat1, np.float64),
            "mask": None if self.mask is None else _host(self.mask, np.uint8),
            "dxf": _host(self.DX, np.float64),
            "kx": _host(self.k[:, None] * np.cos(self.theta)[None, :], np.float64),
            "ky": _host(self.k[:, None] * np.sin(self.theta)[None, :], np.float64),
            "cf": _host(cf, np.int32), "cP": _host(cP, np.int32),
            "clon": _host(clon, np.float64), "clat": _host(clat, np.float64),
            "ccos": _host(np.cos(np.asarray(clat, dtype=np.float64) * np.pi / 180.0), np.float64),
            "cbox": _host(box, np.int32).reshape(-1),
            "fbase": _host(self.iff_wavebounds[:-1], np.int64),
            "fnp": _host(self.NP, np.int32), "fnt": _host(self.nt, np.int32),
            "pptr": _host(np.concatenate([[0], np.cumsum(counts)]), np.int64),
            "plist": _host(pairs[:, 1], np.int32)}
        p = BmParams()
        p.ny, p.nx, p.nf, p.ntheta, p.nc = self.ny, self.nx, self.nf, self.ntheta, self.nc
        for k, v in self._tab.items():
            setattr(p, k, None if v is None else v.ctypes.data)
        if self._h is not None:
            _lib.load().vd_bm_destroy(self._h)
            self._h = None
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.call("vd_bm_create", C.byref(p), C.byref(h))
        self._h = h
        if self._uv is not None:
            au, dyu, bv, dxv = self._uv
            _lib.call("vd_bm_uv_tables", self._h, au.ctypes.data, dyu.ctypes.data, bv.ctypes.data,
                      dxv.ctypes.data)

    # ------------------------------------------------------------- time factors
    def time_factors(self, times):
        """Per frequency the (T, nt_f) factors of the time centres at ``times``
        (basis.py:797-820): window((t - t_c)/tdec) (/ norm_fact) where
        |t - t_c| < tdec, else 0; concatenated as the kernels read them."""
        times = np.atleast_1d(np.asarray(times, dtype=np.float64))
        out = []
        for f in range(self.nf):
            dt = times[:, None] - self.enst[f][None, :]
            g = self.window(dt / self.tdec[f])
            if self.norm_time:
                g = g / self.norm_fact[f]
            g = np.where(np.abs(dt) < self.tdec[f], g, 0.0)
            out.append(g.ravel())
        return np.concatenate(out) if out else np.zeros(0)

    def _workspace(self, T):
        b = C.c_size_t()
        _lib.call("vd_bm_workspace_size", self._h, int(T), C.byref(b))
        return torch.empty(max(1, b.value // 8), dtype=torch.float64, device=self.device)

    # ----------------------------------------------------------------- operators
    def prepare(self, times):
        """Time factors and workspace of a fixed set of basis times, on the
        device, so that ``operg_prepared`` / ``operg_transpose_prepared`` only
        launch kernels (legal inside a captured HIP graph)."""
        times = np.atleast_1d(np.asarray(times, dtype=np.float64))
        return SimpleNamespace(times=times, T=int(times.size),
                               gt=torch.as_tensor(self.time_factors(times), device=self.device),
                               ws=self._workspace(times.size))

    def operg_prepared(self, prep, X, out):
        """out (T, ny, nx) float64 = the fields at the prepared times from the
        float64 device coefficients ``X``."""
        if X.numel() != self.nbasis or out.numel() != prep.T * self.ny * self.nx:
            raise VardynError("operg_prepared: coefficient or output size mismatch")
        _lib.call("vd_bm_operg", self._h, prep.T, ptr(prep.gt), ptr(X), ptr(out), ptr(prep.ws),
                  current_stream_ptr(self.device))
        return out

    def operg_transpose_prepared(self, prep, adssh, adX):
        """adX += sum over the prepared times of operg_transpose(t, adssh[t])."""
        if adX.numel() != self.nbasis or adssh.numel() != prep.T * self.ny * self.nx:
            raise VardynError("operg_transpose_prepared: size mismatch")
        _lib.call("vd_bm_operg_adj", self._h, prep.T, ptr(prep.gt), ptr(adssh), ptr(adX),
                  ptr(prep.ws), current_stream_ptr(self.device))
        return adX

    def operg_batch(self, times, X):
        """SSH fields (T, ny, nx) at ``times`` from coefficients ``X`` (device tensor out)."""
        prep = self.prepare(times)
        x = torch.as_tensor(X, dtype=torch.float64).reshape(-1).to(self.device).contiguous()
        if x.numel() != self.nbasis:
            raise VardynError(f"expected {self.nbasis} coefficients, got {x.numel()}")
        out = torch.empty((prep.T, self.ny, self.nx), dtype=torch.float64, device=self.device)
        return self.operg_prepared(prep, x, out)

    def operg_transpose_batch(self, times, adssh, adX=None):
        """Sum over ``times`` of operg_transpose(t, adssh[t]); ``adssh`` (T, ny, nx)."""
        prep = self.prepare(times)
        a = torch.as_tensor(adssh, dtype=torch.float64).to(self.device).reshape(prep.T, self.ny,
                                                                                self.nx)
        a = torch.nan_to_num(a).contiguous()
        if adX is None:
            adX = torch.zeros(self.nbasis, dtype=torch.float64, device=self.device)
        return self.operg_transpose_prepared(prep, a, adX)

    def ssh2uv_batch(self, ssh):
        """Geostrophic velocities (basis.py:822-833) of SSH fields (T, ny, nx)
        on the device: u (T, ny, nx+1), v (T, ny+1, nx)."""
        if self._uv is None:
            raise VardynError("Basis_bm: built without compute_velocities")
        s = torch.as_tensor(ssh, dtype=torch.float64).to(self.device).reshape(-1, self.ny, self.nx).contiguous()
        T = s.shape[0]
        u = torch.empty((T, self.ny, self.nx + 1), dtype=torch.float64, device=self.device)
        v = torch.empty((T, self.ny + 1, self.nx), dtype=torch.float64, device=self.device)
        _lib.call("vd_bm_ssh2uv", self._h, T, ptr(s), ptr(u), ptr(v), current_stream_ptr(self.device))
        return u, v

    def ssh2uv_adj_batch(self, adu, adv, adssh):
        """adssh (T, ny, nx, device) += the transpose of ssh2uv (basis.py:835-860)
        applied to adu (T, ny, nx+1), adv (T, ny+1, nx)."""
        if self._uv is None:
            raise VardynError("Basis_bm: built without compute_velocities")
        T = adssh.shape[0]
        if adssh.shape != (T, self.ny, self.nx) or not adssh.is_contiguous() or adssh.dtype != torch.float64:
            raise VardynError("ssh2uv_adj_batch: adssh must be a contiguous float64 (T, ny, nx) tensor")
        du = torch.as_tensor(adu, dtype=torch.float64).to(self.device).reshape(T, self.ny, self.nx + 1).contiguous()
        dv = torch.as_tensor(adv, dtype=torch.float64).to(self.device).reshape(T, self.ny + 1, self.nx).contiguous()
        _lib.call("vd_bm_ssh2uv_adj", self._h, T, ptr(du), ptr(dv), ptr(adssh), current_stream_ptr(self.device))
        return adssh

    def operg(self, t, X, State=None):
        """basis.py:863-901: the field at time ``t`` (and with
        ``compute_velocities`` its geostrophic velocities); into
        ``State[name_mod_var]`` (``name_mod_u``, ``name_mod_v``) when a State
        is given, else returned (numpy for numpy ``X``)."""
        ssh = self.operg_batch([t], X)[0]
        out = [ssh]
        if self.compute_velocities:
            u, v = self.ssh2uv_batch(ssh[None])
            out += [u[0], v[0]]
        if State is None:
            if isinstance(X, np.ndarray):
                out = [a.cpu().numpy() for a in out]
            return tuple(out) if self.compute_velocities else out[0]
        names = [self.name_mod_var] + ([self.name_mod_u, self.name_mod_v] if self.compute_velocities else [])
        for n, a in zip(names, out):
            ref = State[n] if n in State else None
            State[n] = a.to(ref.device if torch.is_tensor(ref) else a.device)

    def operg_transpose(self, t, adState):
        """basis.py:903-936: the coefficients' adjoint at time ``t`` (with
        ``compute_velocities`` the velocity adjoints are first added to the
        SSH adjoint); zeroes the adjoint fields like the reference."""
        ad = adState[self.name_mod_var]
        a = torch.as_tensor(ad, dtype=torch.float64).to(self.device).reshape(1, self.ny, self.nx).clone()
        if self.compute_velocities:
            du, dv = adState.get(self.name_mod_u), adState.get(self.name_mod_v)
            if du is not None and dv is not None:
                self.ssh2uv_adj_batch(du, dv, a)
        adX = self.operg_transpose_batch([t], a)
        def zero(x):
            return torch.zeros_like(x) if torch.is_tensor(x) else np.zeros_like(x)
        adState[self.name_mod_var] = zero(ad)
        if self.compute_velocities:
            for n in (self.name_mod_u, self.name_mod_v):
                if adState.get(n) is not None:
                    adState[n] = zero(adState[n])
        return adX.cpu().numpy()
