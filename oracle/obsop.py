"""Restatement of the reference L3 observation operator -- TEST INFRASTRUCTURE ONLY.

Follows ``mapping/src/obsop.py`` (class ``Obsop_interp_l3_jax``) and
``mapping/src/grid.py`` of datlas-ocean/VarDyn:

* ``geo2cart``              -- grid.py:61-98 (WGS84 geodetic -> cartesian, km)
* ``match_obs_times``       -- obsop.py:387-403 (model step gets the nearest
  observation date when it lies within dt/2; first wins on ties)
* ``sparse_op``             -- obsop.py:417-460 (KDTree k=Npix neighbours,
  distance cut at dmax, Gaussian weights normalised per observation)
* ``ObsopOracle.misfit``    -- obsop.py:481-483, 583-628 (padded per-time
  tables; ``(sum w x[idx] - y) / sigma`` with NaN -> 0)
* ``ObsopOracle.adj``       -- obsop.py:630-678 (vjp of the misfit: a
  scatter ``H^T (misfit / sigma)``)

Pinning: ``scipy.spatial.KDTree`` is the library the reference calls; the
selected indices and weights are compared with a brute-force distance sort
over every grid point on small cases (the reference's tests hold no
fixtures for this operator).  ``geo2cart`` is pinned by the WGS84 radii at
the equator and the pole, the time matching by hand-worked cases.  ``adj``
is pinned by the exact identity <H x, y> = <x, H^T y> and by finite
differences of ``misfit``.
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import KDTree


def geo2cart(coords):
    """WGS84 geodetic (lon, lat in degrees) to cartesian km (grid.py:61-98)."""
    A = 6378.137
    E2 = 6.69437999014e-3
    coords = np.asarray(coords).astype(float)
    if coords.ndim == 1:
        coords = np.array([coords])
    lat_rad = np.radians(coords[:, 1])
    lon_rad = np.radians(coords[:, 0])
    r_n = A / (np.sqrt(1 - E2 * (np.sin(lat_rad) ** 2)))
    x = r_n * np.cos(lat_rad) * np.cos(lon_rad)
    y = r_n * np.cos(lat_rad) * np.sin(lon_rad)
    z = r_n * (1 - E2) * np.sin(lat_rad)
    return np.column_stack((x, y, z))


def match_obs_times(T, dt, obs_dates):
    """Model times that receive an observation date (obsop.py:387-403).

    ``T``: model times in seconds; ``obs_dates``: observation dates in seconds,
    in dictionary order.  Returns (t_obs, date_obs) lists.
    """
    t_obs, date_obs = [], []
    dates = list(obs_dates)
    for t in T:
        delta = [(t - d) for d in dates]
        if len(delta) > 0 and np.min(np.abs(delta)) <= dt / 2:
            ind = int(np.argmin(np.abs(delta)))
            if t not in t_obs:
                date_obs.append(dates[ind])
                t_obs.append(t)
    return t_obs, date_obs


def sparse_op(coords_car_grid, lon_obs, lat_obs, Npix, dmax):
    """KDTree interpolation weights (obsop.py:417-460).

    Returns ``(data, row, col)``; empty arrays where the reference's
    ``try/except`` yields an empty operator (no valid neighbour, or Npix == 1
    where ``KDTree.query`` squeezes the neighbour axis).
    """
    coords_car_obs = geo2cart(np.column_stack((lon_obs, lat_obs)))
    tree = KDTree(coords_car_grid)
    D, ind_closest = tree.query(coords_car_obs, k=Npix)
    valid_mask = D <= dmax
    if not np.any(valid_mask) or D.ndim != 2:
        return np.zeros(0), np.zeros(0, dtype=np.int64), np.zeros(0, dtype=np.int64)
    if Npix > 1:
        weights = np.exp(-D ** 2 / (2 * (0.5 * dmax) ** 2)) * valid_mask
        sum_weights = np.sum(weights, axis=1, keepdims=True)
        weights = np.divide(weights, sum_weights, out=np.zeros_like(weights), where=sum_weights > 0)
    else:
        weights = np.ones_like(D)
    row, colk = np.where(valid_mask)
    return weights[row, colk], row.astype(np.int64), ind_closest[row, colk].astype(np.int64)


class ObsopOracle:
    """Padded per-time tables and misfit/adjoint of ``Obsop_interp_l3_jax``.

    ``obs``: dict date_seconds -> (lon, lat, y, err) arrays, in date order.
    """

    def __init__(self, lon_grid, lat_grid, DX, DY, T, dt, obs, Npix=4):
        self.shape = lon_grid.shape
        self.coords_car = geo2cart(np.column_stack((lon_grid.ravel(), lat_grid.ravel())))
        self.Npix = Npix
        self.dmax = Npix * np.mean(np.sqrt(DX ** 2 + DY ** 2)) * 1e-3 * np.sqrt(2) / 2
        self.t_obs, self.date_obs = match_obs_times(T, dt, obs.keys())
        self.obs = obs
        data, rows, cols, ys, errs = {}, {}, {}, {}, {}
        for d, t in zip(self.date_obs, self.t_obs):
            lon, lat, y, err = obs[d]
            data[t], rows[t], cols[t] = sparse_op(self.coords_car, lon, lat, Npix, self.dmax)
            ys[t], errs[t] = np.asarray(y, float), np.asarray(err, float)
        n_data = np.array([data[t].size for t in self.t_obs])
        n_obs = np.array([ys[t].size for t in self.t_obs])
        self.n_data, self.n_obs = n_data, n_obs
        nd, no = (n_data.max(), n_obs.max()) if len(self.t_obs) else (0, 0)
        nt = len(self.t_obs)
        self.data_arr = np.zeros((nt, nd))
        self.row_arr = np.zeros((nt, nd), dtype=np.int64)
        self.col_arr = np.zeros((nt, nd), dtype=np.int64)
        self.y_arr = np.zeros((nt, no))
        self.err_arr = np.ones((nt, no)) * 1e7
        for i, t in enumerate(self.t_obs):
            self.data_arr[i, :n_data[i]] = data[t]
            self.row_arr[i, :n_data[i]] = rows[t]
            self.col_arr[i, :n_data[i]] = cols[t]
            self.y_arr[i, :n_obs[i]] = ys[t]
            self.err_arr[i, :n_obs[i]] = errs[t]

    def is_obs(self, t):
        """Exact membership test used for checkpoint placement (obsop.py:145-146)."""
        return t in self.date_obs

    def is_obs_time(self, t):
        return bool(np.any(np.isclose(t, np.asarray(self.t_obs, dtype=float)))) if self.t_obs else False

    def _idx(self, t):
        return int(np.where(np.asarray(self.t_obs) == t)[0][0])

    def misfit(self, t, x):
        i = self._idx(t)
        x = np.asarray(x, dtype=np.float64).ravel()
        hx = np.zeros(self.y_arr.shape[1])
        with np.errstate(invalid="ignore"):
            np.add.at(hx, self.row_arr[i], self.data_arr[i] * x[self.col_arr[i]])
        m = hx - self.y_arr[i]
        inverr = 1 / self.err_arr[i]
        m = np.where(np.isnan(m), 0, m)
        inverr = np.where(np.isnan(inverr), 0, inverr)
        return inverr * m

    def adj(self, t, x, cot):
        """vjp of ``misfit`` at state ``x`` applied to ``cot`` (obsop.py:630-678)."""
        i = self._idx(t)
        x = np.asarray(x, dtype=np.float64).ravel()
        hx = np.zeros(self.y_arr.shape[1])
        with np.errstate(invalid="ignore"):
            np.add.at(hx, self.row_arr[i], self.data_arr[i] * x[self.col_arr[i]])
        m = hx - self.y_arr[i]
        g = np.where(np.isnan(m), 0, cot / self.err_arr[i])
        adx = np.zeros(x.size)
        np.add.at(adx, self.col_arr[i], self.data_arr[i] * g[self.row_arr[i]])
        return adx.reshape(self.shape)
