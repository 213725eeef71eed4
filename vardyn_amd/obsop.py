"""L3 observation operator with device-resident tables.

Reference: ``Obsop_interp_l3_jax`` (mapping/src/obsop.py:375-678) and
``geo2cart`` (mapping/src/grid.py:61-98), datlas-ocean/VarDyn.

Construction runs on the device (``vd_obs_grid_*``): the grid's cartesian
points are binned into cubes of side >= dmax once, and each observation
searches the 27 cubes around it for its Npix nearest points within dmax,
with the distances rounded as the reference's ``scipy.spatial.KDTree``
rounds them, so the selected grid indices are the KDTree's:

* time matching (obsop.py:387-403): model time t takes the nearest
  observation date when |t - date| <= dt/2 (first one on ties);
* ``_sparse_op`` (obsop.py:417-460): k = Npix neighbours in WGS84 cartesian
  km, kept if D <= dmax, Gaussian weights exp(-D^2 / (2 (dmax/2)^2))
  normalised per observation; the reference's Npix == 1 path (squeezed
  KDTree output) and the no-valid-neighbour path give an empty operator;
* padding (obsop.py:583-604): when a time has fewer entries than the
  largest table, the reference's padded entries (row 0, col 0, weight 0)
  add ``0 * x[0]`` to observation 0; one such entry is kept here so a NaN
  at grid point 0 propagates the same way.

Each time's table is uploaded as CSR by observation (gather for the misfit)
and CSR by grid point (gather for the adjoint, so the scatter H^T is
deterministic, with no atomics).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch
from . import _lib
from ._lib import ObsTable, VardynError
from .qg import current_stream_ptr, ptr


def geo2cart(coords):
    """WGS84 geodetic (lon, lat degrees) to cartesian km (grid.py:61-98)."""
    a_km = 6378.137
    e2 = 6.69437999014e-3
    coords = np.asarray(coords).astype(float)
    if coords.ndim == 1:
        coords = np.array([coords])
    lat = np.radians(coords[:, 1])
    lon = np.radians(coords[:, 0])
    rn = a_km / np.sqrt(1 - e2 * np.sin(lat) ** 2)
    return np.column_stack((rn * np.cos(lat) * np.cos(lon), rn * np.cos(lat) * np.sin(lon),
                            rn * (1 - e2) * np.sin(lat)))


def match_times(T, dt, dates):
    """(t_obs, date_obs) pairs of obsop.py:387-403."""
    dates = list(dates)
    t_obs, date_obs = [], []
    if not dates:
        return t_obs, date_obs
    darr = np.asarray(dates, dtype=np.float64)
    for t in T:
        d = np.abs(t - darr)
        if d.min() <= dt / 2:
            k = int(np.argmin(d))
            if t not in t_obs:
                t_obs.append(float(t))
                date_obs.append(dates[k])
    return t_obs, date_obs


class ObsGrid:
    """The model grid's cartesian points (n x 3 km) binned on the device for
    the neighbour queries of the operator construction (``vd_obs_grid_*``;
    the reference builds a scipy KDTree over the same points,
    obsop.py:424-425)."""

    def __init__(self, coords_car, dmax, device):
        xyz = np.ascontiguousarray(coords_car, dtype=np.float64)
        if xyz.ndim != 2 or xyz.shape[1] != 3:
            raise VardynError("ObsGrid: coordinates must be n x 3")
        self.device = torch.device(device)
        self.dmax = float(dmax)
        self._lib = _lib.load()
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self._lib.vd_obs_grid_create(xyz.ctypes.data, xyz.shape[0], self.dmax, C.byref(h)),
                       "vd_obs_grid_create")
        self._h = h

    def query(self, xyz, npix):
        """(col, w, dist) device tensors of n_obs x npix for the device
        coordinates ``xyz`` (n_obs x 3 km): the npix nearest grid points
        within dmax, nearest first, col -1 / w 0 / dist inf past them."""
        xyz = torch.as_tensor(xyz, dtype=torch.float64, device=self.device).contiguous()
        n = xyz.shape[0]
        col = torch.empty((n, npix), dtype=torch.int32, device=self.device)
        w = torch.empty((n, npix), dtype=torch.float64, device=self.device)
        dist = torch.empty((n, npix), dtype=torch.float64, device=self.device)
        with torch.cuda.device(self.device):
            _lib.check(self._lib.vd_obs_grid_query(self._h, ptr(xyz), n, int(npix), ptr(col), ptr(w), ptr(dist),
                                                   current_stream_ptr()), "vd_obs_grid_query")
        return col, w, dist

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.vd_obs_grid_destroy(h)
            self._h = None


def interp_weights_many(grid, sets, npix):
    """(w, row, col) of ``_sparse_op`` (obsop.py:417-460) for each (lon, lat)
    set of ``sets``, from one device neighbour search of ``grid`` (an
    ``ObsGrid``) over all of them; entries in row order, nearest first within
    a row (the reference's ``np.where(valid_mask)`` order).  Npix == 1, where
    the reference's squeezed KDTree output makes its ``try`` fail, and no
    valid neighbour give an empty operator; non-finite coordinates raise, as
    the reference's KDTree query does."""
    def empty():
        return np.zeros(0), np.zeros(0, np.int64), np.zeros(0, np.int64)
    sizes = [len(lon) for lon, _ in sets]
    if npix == 1 or sum(sizes) == 0:
        return [empty() for _ in sets]
    pts = np.concatenate([geo2cart(np.column_stack((lon, lat))) for lon, lat in sets if len(lon)])
    if not np.all(np.isfinite(pts)):
        raise VardynError("observation coordinates must be finite")
    col, w, _ = grid.query(pts, npix)
    col, w = col.cpu().numpy(), w.cpu().numpy()
    out, start = [], 0
    for n in sizes:
        c, ww = col[start:start + n], w[start:start + n]
        start += n
        valid = c >= 0
        if not np.any(valid):
            out.append(empty())
            continue
        row, k = np.where(valid)
        out.append((ww[row, k], row.astype(np.int64), c[row, k].astype(np.int64)))
    return out


def interp_weights(grid, lon, lat, npix):
    """``interp_weights_many`` for one set of observations."""
    return interp_weights_many(grid, [(lon, lat)], npix)[0]


class _DevTable:
    """One observation time on the device."""

    def __init__(self, n_obs, w, row, col, y, err, ngrid, device, partial):
        if col.size and (col.min() < 0 or col.max() >= ngrid):
            raise VardynError("observation operator column outside the grid")
        order = np.argsort(row, kind="stable")
        w, row, col = w[order], row[order], col[order]
        row_ptr = np.zeros(n_obs + 1, np.int64)
        np.add.at(row_ptr, row + 1, 1)
        row_ptr = np.cumsum(row_ptr)
        t_order = np.argsort(col, kind="stable")
        tcol_all = col[t_order]
        tcol, starts = np.unique(tcol_all, return_index=True)
        tptr = np.append(starts, tcol_all.size)

        def dev(a, dt):
            return torch.as_tensor(np.ascontiguousarray(a, dtype=dt), device=device)

        self.n_obs, self.nnz, self.n_cols = int(n_obs), int(w.size), int(tcol.size)
        self.row_ptr = dev(row_ptr, np.int32)
        self.col = dev(col, np.int32)
        self.w = dev(w, np.float64)
        self.y = dev(y, np.float64)
        self.err = dev(err, np.float64)
        self.tcol = dev(tcol, np.int32)
        self.tptr = dev(tptr, np.int32)
        self.trow = dev(row[t_order], np.int32)
        self.tw = dev(w[t_order], np.float64)
        self.misfit = torch.zeros(max(n_obs, 1), dtype=torch.float64, device=device)
        self.valid = torch.zeros(max(n_obs, 1), dtype=torch.uint8, device=device)
        self.host_y = np.asarray(y, np.float64)
        self.host_err = np.asarray(err, np.float64)
        s = ObsTable()
        s.n_obs, s.nnz, s.n_cols = self.n_obs, self.nnz, self.n_cols
        for f in ("row_ptr", "col", "w", "y", "err", "tcol", "tptr", "trow", "tw"):
            setattr(s, f, getattr(self, f).data_ptr())
        s.partial = partial.data_ptr()
        self.struct = s


class Obsop_interp_l3:
    """Observation operator on SSH with the method names of the reference class.

    ``dict_obs`` maps an observation date (seconds from the window start) to
    ``(lon, lat, y, err)`` arrays (the in-memory form of what the reference
    reads from along-track files, mapping/src/obs.py).
    """

    def __init__(self, lon, lat, DX, DY, T, dt, dict_obs, Npix=4, device=None):
        self.shape = np.shape(lon)
        self.ngrid = int(np.prod(self.shape))
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None \
            else torch.device(device)
        self.coords_car = geo2cart(np.column_stack((np.ravel(lon), np.ravel(lat))))
        self.Npix = int(Npix)
        # obsop.py:405
        self.dmax = self.Npix * np.mean(np.sqrt(np.asarray(DX) ** 2 + np.asarray(DY) ** 2)) \
            * 1e-3 * np.sqrt(2) / 2
        self.dict_obs = dict_obs
        self.dt = float(dt)
        self.T = np.asarray(T, dtype=np.float64)
        self.t_obs, self.date_obs = match_times(self.T, self.dt, dict_obs.keys())
        self._index = {t: i for i, t in enumerate(self.t_obs)}
        self.tables = []
        self._partial = torch.zeros(_lib.VD_REDUCE_SCRATCH, dtype=torch.float64,
                                    device=self.device)

    # ------------------------------------------------------------------
    def process_obs(self, var_bc=None):
        """Build and upload the per-time tables (obsop.py:486-604)."""
        if var_bc is not None:
            raise NotImplementedError("boundary-field anomalies of observations are not supported")
        # a non-finite dmax (NaN metrics) keeps no neighbour, as D <= NaN
        grid = ObsGrid(self.coords_car, self.dmax, self.device) if np.isfinite(self.dmax) and self.dmax > 0 else None
        data = [tuple(np.asarray(a, dtype=np.float64).ravel() for a in self.dict_obs[date])
                for date in self.date_obs]
        if grid is None:
            tabs = [(np.zeros(0), np.zeros(0, np.int64), np.zeros(0, np.int64)) for _ in data]
        else:  # every date in one search
            tabs = interp_weights_many(grid, [(d[0], d[1]) for d in data], self.Npix)
        built = [(d[2].size, w, row, col, d[2], d[3]) for d, (w, row, col) in zip(data, tabs)]
        nd_max = max((b[1].size for b in built), default=0)
        self.tables = []
        for n_obs, w, row, col, y, err in built:
            if w.size < nd_max and n_obs > 0:
                # the reference's padded entries: row 0, col 0, weight 0
                w = np.append(w, 0.0)
                row = np.append(row, 0)
                col = np.append(col, 0)
            self.tables.append(_DevTable(n_obs, w, row, col, y, err, self.ngrid, self.device,
                                         self._partial))
        self.n_obs = np.array([tb.n_obs for tb in self.tables], dtype=np.int64)
        return self

    def is_obs(self, t):
        """Exact date membership, used for checkpoint placement (obsop.py:145-146)."""
        return t in self.date_obs

    def is_obs_time(self, t):
        """obsop.py:606-608 (isclose against the matched model times)."""
        return bool(len(self.t_obs)) and bool(np.any(np.isclose(t, np.asarray(self.t_obs))))

    def index(self, t):
        i = self._index.get(float(t))
        if i is None:
            i = int(np.argmin(np.abs(np.asarray(self.t_obs) - t)))
        return i

    @property
    def n_total(self):
        return int(sum(tb.n_obs for tb in self.tables))

    # ---------------------------------------------------- device operations
    def misfit_dev(self, i, x, jo=None, inv_sigma_r2=1.0):
        """misfit of table ``i`` at device field ``x``; adds sum(m^2)*inv_sigma_r2 to jo."""
        tb = self.tables[i]
        prec = _lib.VD_F64 if x.dtype == torch.float64 else _lib.VD_F32
        _lib.call("vd_obs_misfit", C.byref(tb.struct), prec, ptr(x), ptr(tb.misfit),
                  ptr(tb.valid), ptr(jo), float(inv_sigma_r2), current_stream_ptr(self.device))
        return tb.misfit[:tb.n_obs]

    def adj_dev(self, i, cot, adx):
        """adx += H^T(valid * cot / err) for table ``i``."""
        tb = self.tables[i]
        prec = _lib.VD_F64 if adx.dtype == torch.float64 else _lib.VD_F32
        if cot.dtype != torch.float64 or not cot.is_contiguous():
            raise VardynError("cotangent must be a contiguous float64 device tensor")
        _lib.call("vd_obs_adjoint", C.byref(tb.struct), prec, ptr(cot), ptr(tb.valid), ptr(adx),
                  current_stream_ptr(self.device))
        return adx

    # ------------------------------------------- reference-style interface
    def misfit(self, t, State):
        """``Obsop_interp_l3_jax.misfit`` (obsop.py:656-660): returns the misfit vector."""
        x = State.var["ssh"] if hasattr(State, "var") else State
        return self.misfit_dev(self.index(t), x).clone()

    def adj(self, t, adState, State, misfit):
        """``Obsop_interp_l3_jax.adj`` (obsop.py:668-678): adState += H^T(misfit / err)."""
        i = self.index(t)
        x = State.var["ssh"] if hasattr(State, "var") else State
        self.misfit_dev(i, x)  # refresh the NaN pattern of this state
        ad = adState.var["ssh"] if hasattr(adState, "var") else adState
        cot = torch.as_tensor(misfit, dtype=torch.float64, device=self.device).contiguous()
        self.adj_dev(i, cot, ad)
