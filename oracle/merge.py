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

Pinning (tests/test_oracle_merge.py): the window's closed form (0 at the
edges, 1 at the centre); bilinear interpolation reproduces a field linear in
lon and lat exactly; partition of unity (equal subwindow fields merge to the
same field wherever the weight sum is positive); a single subwindow on the
target grid is the identity; NaN at a cell corner reaches exactly the targets
of the cells around it (brute-force cell search); blend weights at the overlap
ends, its middle and for an overlap of zero length.
"""
from __future__ import annotations

import numpy as np
from scipy.interpolate import RegularGridInterpolator


def _interp(lat1, lon1, values, lat_out, lon_out):
    rgi = RegularGridInterpolator((lat1, lon1), values, method="linear", bounds_error=False,
                                  fill_value=np.nan)
    pts = np.column_stack([np.ravel(lat_out), np.ravel(lon_out)])
    return rgi(pts).reshape(np.shape(lat_out))


def hann(n):
    t = np.linspace(0, 2 * np.pi, n)
    return 0.5 * (1.0 - np.cos(t))


def weights_map(lon_out, lat_out, subwindows):
    """``subwindows``: [(lon_1d, lat_1d), ...]; returns (weights, weights_sum)."""
    single = len(subwindows) == 1
    ws = np.zeros(np.shape(lon_out))
    weights = []
    for lon1, lat1 in subwindows:
        if single:
            w = np.ones((lat1.size, lon1.size))
        else:
            w = hann(lat1.size)[:, None] * hann(lon1.size)[None, :]
        wi = _interp(lat1, lon1, w, lat_out, lon_out)
        ind = ~np.isnan(wi)
        weights.append(wi)
        ws[ind] += wi[ind]
    return weights, ws


def merge_date(lon_out, lat_out, subwindows, fields, weights, weights_sum, mask=None):
    out = np.zeros(np.shape(lon_out))
    for (lon1, lat1), f, w in zip(subwindows, fields, weights):
        v = _interp(lat1, lon1, np.asarray(f, dtype=np.float64), lat_out, lon_out)
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
