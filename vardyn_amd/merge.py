"""Merge of subwindow analyses onto the target grid, on the GPU.

Reference: ``compute_weights_map``, ``_build_interpolator``,
``merge_output_date`` and ``merge_time_windows_outputs``
(mapping/src/run_assimilation.py:351-446, 537-607, 732-830),
datlas-ocean/VarDyn.  The reference splits a large domain into overlapping
subwindows and a long period into overlapping time windows, runs one
assimilation per piece and merges the outputs date by date:

* space (``compute_weights_map`` + ``merge_output_date``): every subwindow's
  field is interpolated onto the target grid (scipy ``RegularGridInterpolator``,
  linear, NaN outside the subwindow) and added with the subwindow's weight,
  the raised-cosine window ``0.5 (1 - cos)`` over ``linspace(0, 2 pi, n)`` on
  each axis (ones for a single subwindow) interpolated the same way, divided
  by the sum of the weights; values that interpolate to NaN are skipped;
  masked target points become NaN;
* time (``merge_time_windows_outputs``): a date inside one window takes that
  window's output; a date inside two or more takes ``W1 a + W2 b`` of the first
  two, ``W2 = 0.5 (1 - cos(pi alpha))``, ``alpha`` the clamped position of the
  date in their overlap (0.5 / 0.5 for an overlap of zero length).

Here the outputs come in as device tensors instead of netCDF files, and a
batch of dates is merged in one call.  Scope: regular (``GRID_GEO``)
subwindow grids with ascending axes and irregular (``GRID_CAR``, ``geo_grid``
False) grids through the reference's Delaunay interpolation (the triangles
and barycentric weights found on the host as LinearNDInterpolator finds
them, applied on the GPU), no longitude wrap (the reference's split-longitude
branch is not built);
NaN gaps are not filled first (``kernel`` None; the reference's astropy fill
is not available); H-grid variables (the C-grid averaging of U/V variables
belongs to the SW model).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib
from ._lib import VardynError
from .qg import current_stream_ptr, ptr


class MergeParams(C.Structure):
    _fields_ = [("ny", C.c_int), ("nx", C.c_int), ("lon", C.c_void_p), ("lat", C.c_void_p),
                ("mask", C.c_void_p), ("nsub", C.c_int), ("sny", C.c_void_p), ("snx", C.c_void_p),
                ("slat", C.c_void_p), ("slon", C.c_void_p), ("tri_slot", C.c_void_p),
                ("tri_idx", C.c_void_p), ("tri_w", C.c_void_p)]


def delaunay_tables(lon_in, lat_in, lon_out, lat_out):
    """Triangle of every target point in the Delaunay triangulation of the
    points (lon, lat) of an irregular subwindow, and its barycentric weights,
    as ``LinearNDInterpolator(points, values)(pts_out)`` finds and computes
    them (run_assimilation.py:436-446; scipy's find_simplex with its default
    tolerance, c_i = sum_j T_ij (x_j - r_j), c_2 = (1 - c_0) - c_1).
    Returns (idx (np, 3) int32, -1 outside the hull; w (np, 3) float64)."""
    from scipy.spatial import Delaunay
    points = np.column_stack([np.ravel(lon_in), np.ravel(lat_in)])
    pts = np.column_stack([np.ravel(lon_out), np.ravel(lat_out)])
    tri = Delaunay(points)
    simplex = tri.find_simplex(pts)
    inside = simplex >= 0
    idx = np.full((pts.shape[0], 3), -1, dtype=np.int32)
    w = np.zeros((pts.shape[0], 3))
    sx = simplex[inside]
    idx[inside] = tri.simplices[sx]
    T = tri.transform[sx]
    d0 = pts[inside, 0] - T[:, 2, 0]
    d1 = pts[inside, 1] - T[:, 2, 1]
    c0 = T[:, 0, 0] * d0 + T[:, 0, 1] * d1
    c1 = T[:, 1, 0] * d0 + T[:, 1, 1] * d1
    w[inside, 0] = c0
    w[inside, 1] = c1
    w[inside, 2] = (1.0 - c0) - c1
    return idx, w


def _axes(lon, lat):
    """1-D axes of a subwindow grid: ``lon_in[0, :]`` and ``lat_in[:, 0]`` of
    2-D arrays (run_assimilation.py:432-433), or 1-D arrays as they are."""
    lon, lat = np.asarray(lon, dtype=np.float64), np.asarray(lat, dtype=np.float64)
    if lon.ndim == 2:
        lon = lon[0, :]
    if lat.ndim == 2:
        lat = lat[:, 0]
    return np.ascontiguousarray(lon), np.ascontiguousarray(lat)


class SubwindowMerge:
    """``compute_weights_map`` for a target grid (2-D ``lon``, ``lat``,
    optional boolean ``mask``) and subwindow grids ``[(lon, lat), ...]``; the
    tables and weights live on ``device``.  ``merge`` then plays
    ``merge_output_date`` for a batch of dates."""

    def __init__(self, lon, lat, subwindows, mask=None, device="cuda", geo_grid=None):
        lon = np.ascontiguousarray(np.asarray(lon, dtype=np.float64))
        lat = np.ascontiguousarray(np.asarray(lat, dtype=np.float64))
        if lon.ndim != 2 or lon.shape != lat.shape:
            raise VardynError("the target grid needs 2-D lon and lat of one shape")
        self.ny, self.nx = lon.shape
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None and torch.cuda.is_available():
            # tensors report their device with its index: compare against that
            self.device = torch.device("cuda", torch.cuda.current_device())
        if not subwindows:
            raise VardynError("no subwindow")
        geo = [True] * len(subwindows) if geo_grid is None else [bool(g) for g in geo_grid]
        if len(geo) != len(subwindows):
            raise VardynError("geo_grid needs one flag per subwindow")
        axes, self.shapes, tri_idx, tri_w = [], [], [], []
        tri_slot = np.full(len(subwindows), -1, dtype=np.int32)
        for s, ((lo, la), g) in enumerate(zip(subwindows, geo)):
            if g:
                ax = _axes(lo, la)
                axes.append(ax)
                self.shapes.append((ax[1].size, ax[0].size))
                continue
            # irregular grid (GRID_CAR): Delaunay triangles of its points
            lo2, la2 = np.asarray(lo, dtype=np.float64), np.asarray(la, dtype=np.float64)
            if lo2.ndim != 2 or lo2.shape != la2.shape:
                raise VardynError("an irregular subwindow needs 2-D lon and lat of one shape")
            i_, w_ = delaunay_tables(lo2, la2, lon, lat)
            tri_slot[s] = len(tri_idx)
            tri_idx.append(i_)
            tri_w.append(w_)
            self.shapes.append(lo2.shape)
            # axes are not read for irregular subwindows: any ascending ones
            axes.append((np.arange(lo2.shape[1], dtype=np.float64), np.arange(lo2.shape[0], dtype=np.float64)))
        self.nsub = len(axes)
        self.geo_grid = geo
        sny = np.array([s[0] for s in self.shapes], dtype=np.int32)
        snx = np.array([s[1] for s in self.shapes], dtype=np.int32)
        slat = np.ascontiguousarray(np.concatenate([la for _, la in axes]))
        slon = np.ascontiguousarray(np.concatenate([lo for lo, _ in axes]))
        m = None
        if mask is not None:
            mask = np.asarray(mask)
            if mask.shape != (self.ny, self.nx):
                raise VardynError(f"mask shape {mask.shape}, target grid {(self.ny, self.nx)}")
            if np.any(mask):
                m = np.ascontiguousarray(mask.astype(np.uint8))
        tri = None
        if tri_idx:
            tri = (tri_slot, np.ascontiguousarray(np.concatenate(tri_idx)),
                   np.ascontiguousarray(np.concatenate(tri_w)))
        p = MergeParams(self.ny, self.nx, lon.ctypes.data, lat.ctypes.data,
                        m.ctypes.data if m is not None else None, self.nsub,
                        sny.ctypes.data, snx.ctypes.data, slat.ctypes.data, slon.ctypes.data,
                        *((a.ctypes.data for a in tri) if tri is not None else (None, None, None)))
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.call("vd_merge_create", C.byref(p), C.byref(h))
        self._h = h
        tot = C.c_longlong()
        _lib.call("vd_merge_info", h, C.byref(tot))
        self.sub_points = int(tot.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.load().vd_merge_destroy(h)
            self._h = None

    @property
    def weights_space(self):
        """Interpolated weight of every subwindow on the target grid
        (nsub x ny x nx, NaN outside the subwindow)."""
        w = torch.empty((self.nsub, self.ny, self.nx), dtype=torch.float64, device=self.device)
        _lib.call("vd_merge_weights", self._h, ptr(w), None, current_stream_ptr(self.device))
        return w

    @property
    def weights_space_sum(self):
        ws = torch.empty((self.ny, self.nx), dtype=torch.float64, device=self.device)
        _lib.call("vd_merge_weights", self._h, None, ptr(ws), current_stream_ptr(self.device))
        return ws

    def pack(self, fields):
        """One device buffer holding every subwindow's (T, ny_s, nx_s) fields."""
        if len(fields) != self.nsub:
            raise VardynError(f"expected {self.nsub} subwindow field stacks, got {len(fields)}")
        T = None
        parts = []
        for f, (ny, nx) in zip(fields, self.shapes):
            f = torch.as_tensor(f, dtype=torch.float64, device=self.device)
            if f.dim() == 2:
                f = f.unsqueeze(0)
            if f.shape[1:] != (ny, nx):
                raise VardynError(f"subwindow field shape {tuple(f.shape[1:])}, grid {(ny, nx)}")
            if T is None:
                T = f.shape[0]
            elif f.shape[0] != T:
                raise VardynError("every subwindow needs the same number of dates")
            parts.append(f.reshape(-1))
        return torch.cat(parts), T

    def _check_device_tensor(self, t, what, shape=None, numel=None):
        if not isinstance(t, torch.Tensor):
            raise VardynError(f"{what} must be a torch tensor on {self.device}")
        if t.device != self.device or t.dtype != torch.float64 or not t.is_contiguous():
            raise VardynError(f"{what} must be a contiguous float64 tensor on {self.device}")
        if shape is not None and tuple(t.shape) != tuple(shape):
            raise VardynError(f"{what} has shape {tuple(t.shape)}, expected {tuple(shape)}")
        if numel is not None and t.numel() != numel:
            raise VardynError(f"{what} holds {t.numel()} values, expected {numel}")

    def merge_packed(self, packed, T, out=None):
        """Merge from a buffer made by ``pack``; ``out`` (T, ny, nx)."""
        T = int(T)
        if T < 1:
            raise VardynError("T must be at least 1")
        self._check_device_tensor(packed, "packed fields", numel=T * self.sub_points)
        if out is None:
            out = torch.empty((T, self.ny, self.nx), dtype=torch.float64, device=self.device)
        self._check_device_tensor(out, "out", shape=(T, self.ny, self.nx))
        _lib.call("vd_merge_apply", self._h, T, ptr(packed), ptr(out), current_stream_ptr(self.device))
        return out

    def merge(self, fields, out=None):
        """``merge_output_date`` for T dates at once: ``fields`` is one
        (T, ny_s, nx_s) (or (ny_s, nx_s)) array per subwindow; returns the
        merged (T, ny, nx) fields."""
        packed, T = self.pack(fields)
        return self.merge_packed(packed, T, out)


def compute_weights_map(lon, lat, subwindows, mask=None, device="cuda", geo_grid=None):
    """run_assimilation.py:351-412: the merger with its weight maps."""
    return SubwindowMerge(lon, lat, subwindows, mask=mask, device=device, geo_grid=geo_grid)


def time_window_weights(date, list_date_start, list_date_end):
    """Which windows a date takes and their weights (run_assimilation.py:785-812):
    ``[]`` outside every window, ``[(i, 1.0)]`` inside one, ``[(i, W1), (j, W2)]``
    for the first two windows that hold it."""
    active = [i for i in range(len(list_date_start)) if list_date_start[i] <= date <= list_date_end[i]]
    if not active:
        return []
    if len(active) == 1:
        return [(active[0], 1.0)]
    i, j = active[0], active[1]
    start, end = list_date_start[j], list_date_end[i]
    dur = end - start
    dur = dur.total_seconds() if hasattr(dur, "total_seconds") else float(dur)
    if dur > 0:
        a = date - start
        a = a.total_seconds() if hasattr(a, "total_seconds") else float(a)
        alpha = min(max(a / dur, 0.0), 1.0)
        w2 = 0.5 * (1.0 - math.cos(math.pi * alpha))
        w1 = 1.0 - w2
    else:
        w1, w2 = 0.5, 0.5
    return [(i, w1), (j, w2)]


def blend(a, b, w1, w2, out=None):
    """``W1 a + W2 b`` of two windows' outputs on the GPU (run_assimilation.py:815-818)."""
    for t, what in ((a, "a"), (b, "b")):
        if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_cuda:
            raise VardynError(f"blend: {what} must be a float64 device tensor")
    if a.shape != b.shape or a.device != b.device:
        raise VardynError("blend needs two tensors of one shape on one device")
    a, b = a.contiguous(), b.contiguous()
    if out is None:
        out = torch.empty_like(a)
    if (not isinstance(out, torch.Tensor) or out.shape != a.shape or out.dtype != torch.float64
            or out.device != a.device or not out.is_contiguous()):
        raise VardynError("blend: out must be a contiguous float64 tensor of the inputs' shape "
                          "and device")
    _lib.call("vd_merge_blend", C.c_longlong(a.numel()), float(w1), float(w2), ptr(a), ptr(b), ptr(out),
              current_stream_ptr(a.device))
    return out
