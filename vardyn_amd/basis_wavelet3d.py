"""The 3-D wavelet basis (``Basis_wavelet3d``, mapping/src/basis.py:2640-2985,
datlas-ocean/VarDyn) with its synthesis and transpose on the GPU.

A reduced control of space-time wave packets: for every space
pseudo-frequency f (log-spaced between 1/lmin and 1/lmax with ratio
1 + facpsp/npsp), direction theta and space centre P (latitude rows
DX_f/facnst km apart around the domain's middle, longitudes the same distance
apart on each row, centres over a fully masked (1/(2f))-box dropped) there is
a cosine and a sine wave sqrt(2) w(dx/DX_f) w(dy/DX_f) cos|sin(k (dx cos theta
+ dy sin theta)) on the centre's box |dx|, |dy| <= DX_f, w(x) = cos^2(pi x/2),
DX_f = npsp/(2 f); each repeats over time blocks (time frequency, time centre
DXt/facnst apart) weighted at time t by the Gaspari-Cohn window
GC(t - t_c, DXt) divided by its Riemann integral (basis.py:2681-2934).  The
coefficient vector is laid out as in the reference: space frequency, time
block, space centre, direction, cos/sin; the prior standard deviation is
sigma_Q for every coefficient.

Here ``set_basis`` follows the reference's construction of the centres and
time blocks and decides, with the reference's box test, which grid points each
centre covers; the library (``vd_w3_*``) then keeps, per 16 x 16-point tile of
the grid, the dense block of space columns of the waves meeting that tile,
and applies the basis and its transpose to batches of basis times as FP64
MFMA products.  ``prepare`` / ``operg_prepared`` / ``operg_transpose_prepared``
are the interface the 4D-Var engines use (kernel launches only, so an
evaluation can be captured in a HIP graph); ``operg`` / ``operg_transpose``
keep the reference's per-time calls.
"""
from __future__ import annotations

import ctypes as C
from types import SimpleNamespace

import numpy as np
import torch

from . import _lib
from ._lib import VardynError
from .qg import current_stream_ptr, ptr

KM2DEG = 1.0 / 110
TILE = 16


class W3Params(C.Structure):
    _fields_ = [("ny", C.c_int), ("nx", C.c_int), ("nf", C.c_int), ("ntheta", C.c_int), ("ntb", C.c_int),
                ("nfam", C.c_int)] + [(n, C.c_void_p) for n in (
                    "lon", "lat", "dxs", "kx", "ky", "fam_f", "fam_lon", "fam_lat", "fam_cos", "fbase",
                    "ffam0", "fnfam", "tslot", "sfam", "sbits")]


def gaspari_cohn(r, c):
    """Gaspari-Cohn window of |r| / c (mapping/src/tools.py:16-47): 1 at 0,
    zero from |r| = c on; vectorised over r."""
    ra = 2 * np.abs(np.asarray(r, dtype=np.float64)) / c
    out = np.zeros_like(ra)
    lo = ra <= 1.0
    hi = (ra > 1.0) & (ra <= 2.0)
    x = ra[lo]
    out[lo] = -0.25 * x ** 5 + 0.5 * x ** 4 + 0.625 * x ** 3 - 5.0 / 3.0 * x ** 2 + 1.0
    x = ra[hi]
    out[hi] = 1.0 / 12.0 * x ** 5 - 0.5 * x ** 4 + 0.625 * x ** 3 + 5.0 / 3.0 * x ** 2 - 5.0 * x + 4.0 \
        - 2.0 / 3.0 / x
    return out


def _frequencies(short, long_, ratio_log):
    """Pseudo-frequencies from 1/short down to 1/long, ascending, periods
    above ``long_`` dropped (basis.py:2698-2714)."""
    f = np.exp(np.arange(np.log(1.0 / short), np.log(1.0 / long_) - ratio_log, -ratio_log)[::-1])
    return f[1.0 / f <= long_]


def _around(mid, low, high, spacing):
    """``spacing``-spaced values from ``mid`` down while above ``low`` and up
    while below ``high`` (the paired aranges of basis.py:2727-2747, 2774-2783)."""
    down = np.arange(mid, low, -spacing)[::-1]
    up = np.arange(mid, high, spacing)[1:]
    return np.concatenate((down, up))


def _host(a, dtype):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


class Basis_wavelet3d:
    """``Basis_wavelet3d`` (basis.py:2640-2985) on the GPU.

    ``State`` needs 2-D ``lon``, ``lat`` (degrees) and an optional boolean
    ``mask`` (True = masked); the keyword arguments are the reference's
    ``config.BASIS`` fields (defaults of BASIS_WAVELET3D,
    mapping/src/config_default.py:1055-1081)."""

    controls_initial_state = False
    time_dependent = True

    def __init__(self, State, facnst=1.0, npsp=3.5, facpsp=1.5, lmin=80.0, lmax=970.0, tmin=2.0, tmax=20.0,
                 sigma_Q=1e-1, name_mod_var="ssh", path_background=None, device=None, **kwargs):
        if path_background is not None:
            raise VardynError("Basis_wavelet3d: a background file (netCDF) is not read here")
        if not torch.cuda.is_available():
            raise VardynError("Basis_wavelet3d needs a GPU (libvardyn_hip)")
        lon = np.asarray(State.lon, dtype=np.float64)
        lat = np.asarray(State.lat, dtype=np.float64)
        if lon.ndim != 2 or lon.shape != lat.shape:
            raise VardynError("Basis_wavelet3d: lon and lat must be 2-D arrays of the grid's shape")
        self.ny, self.nx = lon.shape
        self.shape_phys = (self.ny, self.nx)
        self.nphys = lon.size
        self.lon1d, self.lat1d = lon.ravel().copy(), lat.ravel().copy()
        self.lon_min, self.lon_max = self.lon1d.min(), self.lon1d.max()
        self.lat_min, self.lat_max = self.lat1d.min(), self.lat1d.max()
        m = getattr(State, "mask", None)
        m = None if m is None else np.asarray(m, dtype=bool).ravel()
        self.mask1d = m if (m is not None and np.any(m)) else None
        self.facnst, self.npsp, self.facpsp = float(facnst), float(npsp), float(facpsp)
        self.lmin, self.lmax, self.tmin, self.tmax = float(lmin), float(lmax), float(tmin), float(tmax)
        self.sigma_Q = float(sigma_Q)
        self.name_mod_var = name_mod_var
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None \
            else torch.device(device)
        self._h = None

    @classmethod
    def from_config(cls, config, State, **kw):
        """The reference's ``Basis_wavelet3d(config, State)`` (basis.py:2642-2679)."""
        B = config.BASIS
        names = ("facnst", "npsp", "facpsp", "lmin", "lmax", "tmin", "tmax", "sigma_Q", "name_mod_var",
                 "path_background")
        return cls(State, **{n: getattr(B, n) for n in names if hasattr(B, n)}, **kw)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            try:
                _lib.load().vd_w3_destroy(h)
            except Exception:
                pass
            self._h = None

    # ------------------------------------------------------------------ set-up
    def _box(self, lon_c, lat_c, half):
        """Grid points with |dx|, |dy| <= half (km), dx scaled by cos(lat_c)
        (the reference's point selection, basis.py:2759-2762, 2858-2864)."""
        cl = np.cos(lat_c * np.pi / 180.0)
        return np.where((np.abs((self.lon1d - lon_c) / KM2DEG * cl) <= half)
                        & (np.abs((self.lat1d - lat_c) / KM2DEG) <= half))[0]

    def set_basis(self, time, return_q=False, **kwargs):
        """Centres, time blocks and prior standard deviations
        (basis.py:2681-2848); returns (Xb, Q) with ``return_q``."""
        time = np.asarray(time, dtype=np.float64)
        t_lo, t_hi = time.min(), time.max()
        lon_min, lon_max, lat_min, lat_max = self.lon_min, self.lon_max, self.lat_min, self.lat_max
        if lon_max < lon_min:
            lon_max = lon_max + 360.0
        ratio_log = np.log(1 + self.facpsp / self.npsp)
        fs = _frequencies(self.lmin, self.lmax, ratio_log)
        if fs.size < 2:
            raise VardynError("Basis_wavelet3d: lmin..lmax must give at least two space frequencies")
        ft = _frequencies(self.tmin, self.tmax, ratio_log)
        theta = np.linspace(0, np.pi, int(np.pi * fs[0] / (fs[1] - fs[0]) * self.facpsp))[:-1]
        nfs, nft, ntheta = len(fs), len(ft), len(theta)
        if ntheta < 1 or nft < 1:
            raise VardynError("Basis_wavelet3d: no direction or no time frequency")
        DXs = 1.0 / fs * self.npsp * 0.5
        cen_lon, cen_lat = [], []
        for f in range(nfs):
            rows = _around((lat_min + lat_max) / 2, lat_min - DXs[f] * KM2DEG, lat_max + DXs[f] * KM2DEG,
                           DXs[f] / self.facnst * KM2DEG)
            lo_f, la_f = [], []
            for la in rows:
                cl = np.cos(la * np.pi / 180.0)
                for lo in _around((lon_min + lon_max) / 2, lon_min - DXs[f] / cl * KM2DEG,
                                  lon_max + DXs[f] / cl * KM2DEG, DXs[f] / self.facnst / cl * KM2DEG):
                    if self.mask1d is not None and np.all(self.mask1d[self._box(lo, la, 0.5 / fs[f])]):
                        continue  # a centre over masked points only (basis.py:2749-2768)
                    lo_f.append(lo)
                    la_f.append(la)
            cen_lon.append(np.asarray(lo_f, dtype=np.float64))
            cen_lat.append(np.asarray(la_f, dtype=np.float64))
        NS = np.array([c.size for c in cen_lon], dtype=np.int64)
        DXt = 1.0 / ft * self.npsp * 0.5
        enst, norm = [], []
        for i in range(nft):
            enst.append(_around((t_lo + t_hi) / 2, t_lo - DXt[i], t_hi + DXt[i], DXt[i] / self.facnst))
            tt = np.linspace(-DXt[i], DXt[i])
            cum = np.concatenate([[0.0], np.cumsum(gaspari_cohn(tt[:-1], DXt[i]) * (tt[1:] - tt[:-1]))])
            norm.append(cum.max())
        ntb = int(sum(e.size for e in enst))
        nwf = np.where(NS > 0, 2 * ntb * ntheta * NS, 0)
        self.fs_wavebounds = [0] + [int(v) for v in np.cumsum(nwf)]
        self.nbasis = int(self.fs_wavebounds[-1])
        self.fs, self.ft, self.theta, self.nfs, self.nft, self.ntheta = fs, ft, theta, nfs, nft, ntheta
        self.k = 2 * np.pi * fs
        self.DXs, self.DXt, self.NS, self.ntb = DXs, DXt, NS, ntb
        self.ENSLON, self.ENSLAT, self.ENST, self.norm_fact = cen_lon, cen_lat, enst, norm
        self.time = time
        self._build_tables()
        if return_q:
            return np.zeros(self.nbasis), np.full(self.nbasis, self.sigma_Q)

    def _build_tables(self):
        ntx, nty = -(-self.nx // TILE), -(-self.ny // TILE)
        fam_f, fam_lon, fam_lat = [], [], []
        slot_tile, slot_fam, slot_bits = [], [], []
        fam = 0
        nnz = 0
        for f in range(self.nfs):
            for P in range(int(self.NS[f])):
                lo, la = self.ENSLON[f][P], self.ENSLAT[f][P]
                pts = self._box(lo, la, self.DXs[f])
                if self.mask1d is not None:
                    pts = pts[~self.mask1d[pts]]
                nnz += pts.size
                y, x = pts // self.nx, pts % self.nx
                tiles = (y // TILE) * ntx + x // TILE
                local = (y % TILE) * TILE + x % TILE
                ut, inv = np.unique(tiles, return_inverse=True)
                bits = np.zeros((ut.size, 4), dtype=np.uint64)
                np.bitwise_or.at(bits, (inv, local // 64), np.left_shift(np.uint64(1), (local % 64).astype(np.uint64)))
                slot_tile.append(ut)
                slot_fam.append(np.full(ut.size, fam, dtype=np.int64))
                slot_bits.append(bits)
                fam_f.append(f)
                fam_lon.append(lo)
                fam_lat.append(la)
                fam += 1
        self.nfam = fam
        self.nnz_space = nnz * 2 * self.ntheta  # entries of the reference's sparse space matrices
        st = np.concatenate(slot_tile) if slot_tile else np.zeros(0, dtype=np.int64)
        sf = np.concatenate(slot_fam) if slot_fam else np.zeros(0, dtype=np.int64)
        sb = np.concatenate(slot_bits) if slot_bits else np.zeros((0, 4), dtype=np.uint64)
        order = np.lexsort((sf, st))  # by tile, families ascending inside a tile
        st, sf, sb = st[order], sf[order], sb[order]
        tslot = np.concatenate([[0], np.cumsum(np.bincount(st, minlength=ntx * nty))])
        ffam0 = np.concatenate([[0], np.cumsum(self.NS)[:-1]])
        self._tab = {
            "lon": _host(self.lon1d, np.float64), "lat": _host(self.lat1d, np.float64),
            "dxs": _host(self.DXs, np.float64),
            "kx": _host(self.k[:, None] * np.cos(self.theta)[None, :], np.float64),
            "ky": _host(self.k[:, None] * np.sin(self.theta)[None, :], np.float64),
            "fam_f": _host(fam_f, np.int32), "fam_lon": _host(fam_lon, np.float64),
            "fam_lat": _host(fam_lat, np.float64),
            "fam_cos": _host(np.cos(np.asarray(fam_lat, dtype=np.float64) * np.pi / 180.0), np.float64),
            "fbase": _host(self.fs_wavebounds[:-1], np.int64), "ffam0": _host(ffam0, np.int32),
            "fnfam": _host(self.NS, np.int32), "tslot": _host(tslot, np.int64), "sfam": _host(sf, np.int32),
            "sbits": _host(sb.reshape(-1), np.uint64)}
        p = W3Params()
        p.ny, p.nx, p.nf, p.ntheta, p.ntb, p.nfam = self.ny, self.nx, self.nfs, self.ntheta, self.ntb, self.nfam
        for k, v in self._tab.items():
            setattr(p, k, v.ctypes.data if v.size else None)
        if self._h is not None:
            _lib.load().vd_w3_destroy(self._h)
            self._h = None
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.call("vd_w3_create", C.byref(p), C.byref(h))
        self._h = h
        nb, ne, nw = C.c_longlong(), C.c_longlong(), C.c_longlong()
        _lib.call("vd_w3_info", h, C.byref(nb), C.byref(ne), C.byref(nw))
        if nb.value != self.nbasis:
            raise VardynError("Basis_wavelet3d: coefficient count differs between host and library")
        self.nentries, self.nwave = ne.value, nw.value
        self.block_bytes = self.nentries * TILE * TILE * 8  # the tile blocks of the space matrix

    # ------------------------------------------------------------- time factors
    def time_factors(self, times):
        """(T, nf, ntb) factors of the time blocks at ``times`` (days):
        GC(t - t_c, DXt)/norm where |t - t_c| < DXt, else 0; the same for
        every space frequency (basis.py:2912-2934)."""
        times = np.atleast_1d(np.asarray(times, dtype=np.float64))
        cols = []
        for i in range(self.nft):
            dt = times[:, None] - self.ENST[i][None, :]
            g = gaspari_cohn(dt, self.DXt[i]) / self.norm_fact[i]
            cols.append(np.where(np.abs(dt) < self.DXt[i], g, 0.0))
        g = np.concatenate(cols, axis=1)  # (T, ntb)
        return np.ascontiguousarray(np.broadcast_to(g[:, None, :], (times.size, self.nfs, self.ntb)))

    def _workspace(self):
        b = C.c_size_t()
        _lib.call("vd_w3_workspace_size", self._h, C.byref(b))
        return torch.empty(max(1, b.value // 8), dtype=torch.float64, device=self.device)

    # ----------------------------------------------------------------- operators
    def prepare(self, times):
        """Time factors and workspace of a fixed set of basis times, on the
        device, so that ``operg_prepared`` / ``operg_transpose_prepared`` only
        launch kernels (legal inside a captured HIP graph)."""
        times = np.atleast_1d(np.asarray(times, dtype=np.float64))
        return SimpleNamespace(times=times, T=int(times.size),
                               gt=torch.as_tensor(self.time_factors(times), device=self.device),
                               ws=self._workspace())

    def operg_prepared(self, prep, X, out):
        """out (T, ny, nx) float64 = the fields at the prepared times from the
        float64 device coefficients ``X``."""
        if X.numel() != self.nbasis or out.numel() != prep.T * self.ny * self.nx:
            raise VardynError("operg_prepared: coefficient or output size mismatch")
        _lib.call("vd_w3_operg", self._h, prep.T, ptr(prep.gt), ptr(X), ptr(out), ptr(prep.ws),
                  current_stream_ptr(self.device))
        return out

    def operg_transpose_prepared(self, prep, ad, adX):
        """adX += sum over the prepared times of operg_transpose(t, ad[t])."""
        if adX.numel() != self.nbasis or ad.numel() != prep.T * self.ny * self.nx:
            raise VardynError("operg_transpose_prepared: size mismatch")
        _lib.call("vd_w3_operg_adj", self._h, prep.T, ptr(prep.gt), ptr(ad), ptr(adX), ptr(prep.ws),
                  current_stream_ptr(self.device))
        return adX

    def _coefficients(self, X):
        x = torch.as_tensor(X, dtype=torch.float64).reshape(-1).to(self.device).contiguous()
        if x.numel() != self.nbasis:
            raise VardynError(f"expected {self.nbasis} coefficients, got {x.numel()}")
        return x

    def operg_batch(self, times, X):
        """Fields (T, ny, nx) at ``times`` from coefficients ``X`` (device tensor out)."""
        prep = self.prepare(times)
        out = torch.empty((prep.T, self.ny, self.nx), dtype=torch.float64, device=self.device)
        return self.operg_prepared(prep, self._coefficients(X), out)

    def operg_transpose_batch(self, times, ad, adX=None):
        """Sum over ``times`` of operg_transpose(t, ad[t]); ``ad`` (T, ny, nx)."""
        prep = self.prepare(times)
        a = torch.as_tensor(ad, dtype=torch.float64).to(self.device).reshape(prep.T, self.ny, self.nx)
        a = torch.nan_to_num(a).contiguous()
        if adX is None:
            adX = torch.zeros(self.nbasis, dtype=torch.float64, device=self.device)
        return self.operg_transpose_prepared(prep, a, adX)

    def operg(self, t, X, State=None):
        """basis.py:2936-2960: the field at time ``t`` (days); into
        ``State[name_mod_var]`` when a State is given, else returned (numpy
        for numpy ``X``)."""
        phi = self.operg_batch([t], X)[0]
        if State is None:
            return phi.cpu().numpy() if isinstance(X, np.ndarray) else phi
        ref = State[self.name_mod_var]
        State[self.name_mod_var] = phi.to(ref.device if torch.is_tensor(ref) else phi.device)

    def operg_transpose(self, t, adState):
        """basis.py:2962-2985: the coefficients' adjoint at time ``t``;
        zeroes ``adState[name_mod_var]`` like the reference."""
        ad = adState[self.name_mod_var]
        if ad is None:
            ad = np.zeros(self.shape_phys)
        adX = self.operg_transpose_batch([t], torch.as_tensor(ad).reshape(1, self.ny, self.nx))
        adState[self.name_mod_var] = torch.zeros_like(torch.as_tensor(ad)) if torch.is_tensor(ad) \
            else np.zeros_like(ad)
        return adX.cpu().numpy()
