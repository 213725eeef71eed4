"""FP64 restatement of the subwindow and time-window merges.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``mapping/src/run_assimilation.py`` of datlas-ocean/VarDyn:

* ``weights_map`` -- ``compute_weights_map`` (:351-412) with the regular-grid
  branch of ``_build_interpolator`` (:414-446): per subwindow the window
  ``winy[:, None] * winx[None, :]``, ``win = 0.5 (1 - cos(linspace(0, 2 pi, n)))``
  (ones for a single subwindow), interpolated onto the target points with
  ``RegularGridInterpolator((lat_1d, lon_1d), ..., method='linear',
  bounds_error=False, fill_value=nan)``; the sum of the non-NaN weights.
* ``merge_date`` -- ``merge_output_date`` (:537-607) for one date without the
  files and with ``kernel`` None: every subwindow field interpolated the same
  way, ``dict_var[ind] += (w * v / wsum)[ind]`` over ``ind = ~isnan(v)`` in
  subwindow order, NaN on masked points.
* ``time_blend`` -- the weights of ``merge_time_windows_outputs`` (:785-818).
* the irregular-grid branch of ``_build_interpolator`` (:436-446): a
  subwindow given by 2-D ``lon``, ``lat`` arrays with ``geo`` False is
  interpolated with ``LinearNDInterpolator`` over the points ``(lon, lat)``
  (its Delaunay triangulation, linear on each triangle, NaN outside the
  convex hull), the reference's own scipy call.

Pinning (tests/test_oracle_merge.py): the window's closed form (0 at the
edges, 1 at the centre); bilinear interpolation reproduces a field linear in
lon and lat exactly; partition of unity (equal subwindow fields merge to the
same field wherever the weight sum is positive); a single subwindow on the
target grid is the identity; NaN at a cell corner reaches exactly the targets
of the cells around it (brute-force cell search); blend weights at the overlap
ends, its middle and for an overlap of zero length; for the irregular branch, a
field linear in lon and lat reproduced exactly inside the hull and NaN
exactly at the targets outside it (convex-hull test from the hull's facet
equations).
"""
from __future__ import annotations

import numpy as np
from scipy.interpolate import LinearNDInterpolator, RegularGridInterpolator


def _interp(lat1, lon1, values, lat_out, lon_out):
    rgi = RegularGridInterpolator((lat1, lon1), values, method="linear", bounds_error=False,
                                  fill_value=np.nan)
    pts = np.column_stack([np.ravel(lat_out), np.ravel(lon_out)])
    return rgi(pts).reshape(np.shape(lat_out))


def _interp_tri(lon2, lat2, values, lat_out, lon_out):
    """run_assimilation.py:436-446: LinearNDInterpolator over (lon, lat) points."""
    points = np.column_stack([np.ravel(lon2), np.ravel(lat2)])
    pts_out = np.column_stack([np.ravel(lon_out), np.ravel(lat_out)])
    return LinearNDInterpolator(points, np.ravel(values))(pts_out).reshape(np.shape(lat_out))


def _sub_interp(sub, geo, values, lat_out, lon_out):
    lo, la = sub
    if geo:
        lo = lo[0, :] if np.ndim(lo) == 2 else lo
        la = la[:, 0] if np.ndim(la) == 2 else la
        return _interp(la, lo, values, lat_out, lon_out)
    return _interp_tri(lo, la, values, lat_out, lon_out)


def _sub_shape(sub, geo):
    lo, la = sub
    if not geo:
        return np.shape(lo)
    n1 = np.shape(lo)[-1]
    n0 = np.shape(la)[0]
    return (n0, n1)


def hann(n):
    t = np.linspace(0, 2 * np.pi, n)
    return 0.5 * (1.0 - np.cos(t))


def weights_map(lon_out, lat_out, subwindows, geo=None):
    """``subwindows``: [(lon, lat), ...] (1-D axes, or 2-D arrays; ``geo[s]``
    False: an irregular grid); returns (weights, weights_sum)."""
    geo = [True] * len(subwindows) if geo is None else list(geo)
    single = len(subwindows) == 1
    ws = np.zeros(np.shape(lon_out))
    weights = []
    for sub, g in zip(subwindows, geo):
        ny, nx = _sub_shape(sub, g)
        if single:
            w = np.ones((ny, nx))
        else:
            w = hann(ny)[:, None] * hann(nx)[None, :]
        wi = _sub_interp(sub, g, w, lat_out, lon_out)
        ind = ~np.isnan(wi)
        weights.append(wi)
        ws[ind] += wi[ind]
    return weights, ws


def merge_date(lon_out, lat_out, subwindows, fields, weights, weights_sum, mask=None, geo=None):
    geo = [True] * len(subwindows) if geo is None else list(geo)
    out = np.zeros(np.shape(lon_out))
    for sub, g, f, w in zip(subwindows, geo, fields, weights):
        v = _sub_interp(sub, g, np.asarray(f, dtype=np.float64), lat_out, lon_out)
        ind = ~np.isnan(v)
        with np.errstate(invalid="ignore", divide="ignore"):
            out[ind] += (w * v / weights_sum)[ind]
    if mask is not None and np.any(mask):
        out[np.asarray(mask, dtype=bool)] = np.nan
    return out


def time_blend(date, start_j, end_i):
    """(W1, W2) for a date in the overlap [start_j, end_i] of windows i < j."""
    dur = end_i - start_j
    if dur > 0:
        alpha = min(max((date - start_j) / dur, 0.0), 1.0)
        w2 = 0.5 * (1.0 - np.cos(np.pi * alpha))
        return 1.0 - w2, w2
    return 0.5, 0.5


__all__ = ["hann", "weights_map", "merge_date", "time_blend"]
