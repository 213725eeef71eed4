"""Observation ingestion: along-track (nadir) and swath SSH datasets to the
observation dictionary of the L3 operator.

Reference: ``Obs`` / ``_obs_alti`` (mapping/src/obs.py:20-330) and the merge
of ``Obsop_interp_l3.process_obs`` (mapping/src/obsop.py:486-547),
datlas-ocean/VarDyn.  The reference opens netCDF files with xarray, writes
one netCDF file per assimilation date and lists their paths in ``dict_obs``;
the L3 operator reads them back and merges the datasets of a date.  netCDF
readers of netCDF-4 files are not available here: a dataset comes in as arrays
(its time axis first; ``read_obs_files`` reads netCDF classic files into
them with scipy) and the result is the in-memory dictionary
``{date_seconds: (lon, lat, y, err)}`` that ``vardyn_amd.obsop`` takes --
the same numbers, without the files in between.

Per dataset, as ``_obs_alti`` does it:

1. longitudes to the grid's convention (``0_360`` or ``-180_180``, obs.py:185-189);
2. the box ``bbox = [lon_min, lon_max, lat_min, lat_max]`` (obs.py:191-196):
   samples outside are dropped (a swath keeps its shape, outside pixels
   become NaN, time rows with no pixel inside are dropped);
3. per assimilation date d, the samples with time in [d - dt/2, d + dt/2]
   (label slicing on a sorted time axis, both ends included -- a sample
   exactly on a boundary goes to both dates; [d - dt/2, d + dt/2) when the
   times are not sorted, obs.py:214-225); a date without a finite position
   has no observation (obs.py:230-232);
4. the variable, or a signed sum of several (``combine_var``, obs.py:239-247);
5. plus or minus the MDT interpolated at the samples (obs.py:250-256): linear
   on the MDT's lon/lat grid, extrapolated outside it, as
   ``read_auxdata``'s ``RegularGridInterpolator(..., bounds_error=False,
   fill_value=None)`` (mapping/src/tools.py:204-234);
6. synthetic Gaussian noise of standard deviation ``synthetic_noise``
   (obs.py:258-259; drawn from ``rng`` here, from numpy's global state there);
7. values with |v| > ``varmax`` set to NaN (obs.py:261-262);
8. ``subsampling``: means over consecutive blocks of that many samples along
   every axis, the remainder trimmed, NaN skipped (xarray's
   ``coarsen(..., boundary='trim').mean()``, obs.py:264-268), for the
   positions and times too (obs.py:290-297);
9. the error: an error field interpolated at the samples, times ``facR``
   (obs.py:270-277), else the L3 operator's ``sigma_noise``, else 1
   (obsop.py:535-541).

Then the datasets of a date are flattened and concatenated in dataset order
(obsop.py:513-547).  Samples whose position or value is NaN are dropped: the
L3 operator gives them a zero misfit, so J and its gradient are unchanged.

Differences: with both ``subsampling`` and an error field the reference
interpolates the error at the positions before subsampling and fails on the
shape mismatch; here it is interpolated at the subsampled positions.
"""
from __future__ import annotations

import numpy as np
from scipy.interpolate import RegularGridInterpolator


_TIME_UNITS = {"seconds": 1, "second": 1, "s": 1, "minutes": 60, "minute": 60, "hours": 3600, "hour": 3600,
               "h": 3600, "days": 86400, "day": 86400, "d": 86400}


def _cf_decode(var):
    """CF decoding of a netCDF variable as xarray's default
    ``decode_cf`` does it: ``_FillValue`` / ``missing_value`` to NaN, then
    ``scale_factor`` and ``add_offset``; a ``units`` of "<unit> since <date>"
    gives datetime64[ns] (the offsets rounded to whole nanoseconds).  Values
    are unpacked in FP64 (xarray may unpack small integers with float32
    packing attributes to float32: a difference at float32 rounding)."""
    attrs = var._attributes
    data = np.array(var[:], copy=True)
    units = attrs.get("units", b"")
    units = units.decode() if isinstance(units, bytes) else str(units)
    out = data.astype(np.float64)
    for key in ("_FillValue", "missing_value"):
        if key in attrs:
            fill = np.asarray(attrs[key]).ravel()
            out[np.isin(data, fill)] = np.nan
    if "scale_factor" in attrs:
        out = out * float(np.asarray(attrs["scale_factor"]).ravel()[0])
    if "add_offset" in attrs:
        out = out + float(np.asarray(attrs["add_offset"]).ravel()[0])
    if " since " in units:
        unit, origin = units.split(" since ", 1)
        scale = _TIME_UNITS.get(unit.strip().lower())
        if scale is None:
            raise ValueError(f"unsupported time unit {unit!r}")
        base = np.datetime64(origin.strip().replace(" ", "T").rstrip("Z"), "ns")
        ns = np.round(out * (scale * 1e9))
        t = base + np.where(np.isfinite(ns), ns, 0).astype(np.int64).astype("timedelta64[ns]")
        return np.where(np.isfinite(ns), t, np.datetime64("NaT"))
    return out


def read_obs_files(path, name_time, name_lon, name_lat, names):
    """Read an observation dataset from netCDF classic (version 3) files, the
    file selection of obs.py:120-147: ``path`` itself when it names one
    ``.nc`` file, else the files matching ``path`` (a glob pattern) or
    ``path*.nc``, combined along the time axis in the order of their first
    time (xarray's ``open_mfdataset`` by coordinates).  Variables are
    CF-decoded (``_cf_decode``); those without the time dimension come from
    the first file.  Returns ``{name: array}`` for the time, lon, lat and the
    ``names`` variables, ready for ``obs_alti``.  (The reference opens the
    files with xarray, which also reads netCDF-4 / HDF5 files through the
    netCDF4 library; neither is installed here, and scipy reads only the
    classic format.)"""
    import glob
    from scipy.io import netcdf_file
    if ".nc" in path and "*" not in path:
        files = [path]
    else:
        files = sorted(glob.glob(path if "*" in path else f"{path}*.nc"))
    if not files:
        raise FileNotFoundError(f"no observation file matches {path!r}")
    wanted = [name_time, name_lon, name_lat] + [n for n in names if n not in (name_time, name_lon, name_lat)]
    parts = []
    for f in files:
        with netcdf_file(f, "r", mmap=False) as nc:
            tdim = nc.variables[name_time].dimensions[0]
            parts.append({n: (_cf_decode(nc.variables[n]), nc.variables[n].dimensions[:1] == (tdim,))
                          for n in wanted})
    parts.sort(key=lambda p: p[name_time][0][0] if p[name_time][0].size else np.datetime64("NaT"))
    out = {}
    for n in wanted:
        if parts[0][n][1]:
            out[n] = np.concatenate([p[n][0] for p in parts], axis=0)
        else:
            out[n] = parts[0][n][0]
    return out


def convert_lon(lon, lon_unit):
    """obs.py:185-189 (and tools.py:209-212 for auxiliary grids)."""
    lon = np.asarray(lon, dtype=np.float64)
    finite = lon[np.isfinite(lon)]
    if finite.size == 0:
        return lon
    if np.sign(finite.min()) == -1 and lon_unit == "0_360":
        return lon % 360
    if (np.sign(finite.min()) >= 0 or finite.max() > 180) and lon_unit == "-180_180":
        return (lon + 180) % 360 - 180
    return lon


def grid_interpolator(lon, lat, field, lon_unit="-180_180"):
    """``read_auxdata`` (tools.py:204-234) for a regular grid: 1-D ``lon``,
    ``lat``, ``field`` as (lat, lon) or (lon, lat) (a 3-D field is averaged
    over its first axis); linear, extrapolated outside the grid."""
    lon = convert_lon(lon, lon_unit)
    lat = np.asarray(lat, dtype=np.float64)
    data = np.asarray(field, dtype=np.float64).squeeze()
    if data.ndim == 3:
        data = np.mean(data, 0)
    if data.shape[1] == lon.size:
        data = data.transpose()
    order = np.argsort(lon, kind="stable")
    return RegularGridInterpolator((lon[order], lat), data[order], bounds_error=False,
                                   fill_value=None)


def read_auxdata(file_aux, name_var, lon_unit="-180_180"):
    """``read_auxdata`` (tools.py:204-234) of a netCDF classic file: the
    field named by ``name_var['data']``, ``['mdt']`` or ``['var']`` (that
    order) on the 1-D ``name_var['lon']`` / ``['lat']`` axes, CF-decoded, as
    the ``grid_interpolator`` of it (the MDT and error fields of
    ``obs_alti``)."""
    from scipy.io import netcdf_file
    with netcdf_file(file_aux, "r", mmap=False) as nc:
        key = next(k for k in ("data", "mdt", "var") if k in name_var)
        lon = _cf_decode(nc.variables[name_var["lon"]])
        lat = _cf_decode(nc.variables[name_var["lat"]])
        data = _cf_decode(nc.variables[name_var[key]])
    if lon.ndim != 1 or lat.ndim != 1:
        raise ValueError("read_auxdata: 2-D coordinates (the reference's LinearNDInterpolator branch) are not read")
    return grid_interpolator(lon, lat, data, lon_unit)


def _coarsen_mean(a, s):
    """xarray ``coarsen({dim: s for dim in dims}, boundary='trim').mean()``."""
    a = np.asarray(a, dtype=np.float64)
    shape = tuple(n // s for n in a.shape)
    a = a[tuple(slice(0, n * s) for n in shape)]
    blocks = a.reshape(sum(((n, s) for n in shape), ()))
    axes = tuple(range(1, 2 * len(shape), 2))
    with np.errstate(invalid="ignore"):
        cnt = np.sum(np.isfinite(blocks), axis=axes)
        tot = np.nansum(blocks, axis=axes)
        return np.where(cnt > 0, tot / np.maximum(cnt, 1), np.nan)


def obs_alti(time, lon, lat, var, assim_dates, dt_timestep, bbox=None, lon_unit="-180_180",
             combine_signs=None, mdt=None, add_mdt=False, substract_mdt=False, err_field=None,
             facR=None, synthetic_noise=None, rng=None, varmax=None, subsampling=None):
    """``_obs_alti`` (obs.py:172-330) on arrays.

    ``time``: times of the first axis (numbers in any unit, or datetime64);
    ``lon``, ``lat``: the samples' positions, shaped like ``var`` (or 1-D
    along the time axis for a nadir track); ``var``: one array, or a list of
    arrays combined with ``combine_signs``; ``assim_dates``, ``dt_timestep``:
    the assimilation dates and step in ``time``'s unit; ``mdt`` / ``err_field``:
    callables of (lon, lat) (``grid_interpolator``).  Returns
    ``{date: (lon, lat, var, err or None)}`` for the dates with observations,
    arrays keeping the dataset's shape (NaN where masked).
    """
    time = np.asarray(time)
    lon = convert_lon(lon, lon_unit)
    lat = np.asarray(lat, dtype=np.float64)
    vars_ = var if isinstance(var, (list, tuple)) else [var]
    vars_ = [np.asarray(v, dtype=np.float64) for v in vars_]
    if combine_signs is None:
        combine_signs = [1] * len(vars_)
    if len(vars_) > 1 and len(combine_signs) != len(vars_):
        raise ValueError("one sign per combined variable")
    shape = vars_[0].shape
    lon = np.broadcast_to(lon.reshape(lon.shape + (1,) * (len(shape) - lon.ndim)), shape)
    lat = np.broadcast_to(lat.reshape(lat.shape + (1,) * (len(shape) - lat.ndim)), shape)
    if time.shape[0] != shape[0]:
        raise ValueError("the first axis of the dataset must be the time axis")

    # 2. the box: drop time rows (and swath columns) with no sample inside, mask the rest
    if bbox is not None:
        inside = (bbox[0] <= lon) & (bbox[1] >= lon) & (bbox[2] <= lat) & (bbox[3] >= lat)
        rows = np.any(inside.reshape(shape[0], -1), axis=1)
        sel = [rows]
        for ax in range(1, len(shape)):
            other = tuple(a for a in range(len(shape)) if a != ax)
            sel.append(np.any(inside, axis=other))
        ix = np.ix_(*sel)
        keep = inside[ix]
        time = time[rows]
        lon = np.where(keep, lon[ix], np.nan)
        lat = np.where(keep, lat[ix], np.nan)
        vars_ = [np.where(keep, v[ix], np.nan) for v in vars_]

    sorted_t = time.size < 2 or bool(np.all(time[1:] >= time[:-1]))
    out = {}
    for d in assim_dates:
        t1, t2 = d - dt_timestep / 2, d + dt_timestep / 2
        rows = (time >= t1) & (time <= t2) if sorted_t else (time >= t1) & (time < t2)
        lo, la = lon[rows], lat[rows]
        if lo.size == 0 or not np.any(np.isfinite(lo * la)):
            continue
        v = sum(s * x[rows] for s, x in zip(combine_signs, vars_))
        if mdt is not None and (add_mdt or substract_mdt):
            m = mdt((lo, la))
            v = v + (1 if add_mdt else -1) * m
        if synthetic_noise is not None:
            g = rng if rng is not None else np.random.default_rng()
            v = v + g.normal(0, synthetic_noise, v.size).reshape(v.shape)
        if varmax is not None:
            v = np.where(np.abs(v) <= varmax, v, np.nan)
        if subsampling is not None:
            v = _coarsen_mean(v, subsampling)
            lo = _coarsen_mean(lo, subsampling)
            la = _coarsen_mean(la, subsampling)
        err = None
        if err_field is not None:
            err = err_field((lo, la))
            if facR is not None:
                err = err * facR
        out[d] = (lo, la, v, err)
    return out


def merge_dict_obs(per_dataset, sigma_noise=None, t0=None):
    """The L3 operator's merge of the datasets of each date
    (obsop.py:513-547): flatten, error from the dataset or ``sigma_noise``
    or 1, concatenate in dataset order; NaN samples dropped.  Keys become
    ``date - t0`` (seconds from the window start for ``vardyn_amd.obsop``)."""
    dates = sorted({d for ds in per_dataset for d in ds})
    out = {}
    for d in dates:
        parts = []
        for k, ds in enumerate(per_dataset):
            if d not in ds:
                continue
            lo, la, v, err = (None if a is None else np.asarray(a, dtype=np.float64).ravel()
                              for a in ds[d])
            if err is None:
                s = sigma_noise[k] if isinstance(sigma_noise, (list, tuple)) else sigma_noise
                err = (s if s is not None else 1.0) * np.ones_like(v)
            parts.append((lo, la, v, err))
        lo, la, v, err = (np.concatenate(x) for x in zip(*parts))
        ok = np.isfinite(lo) & np.isfinite(la) & np.isfinite(v)
        key = d if t0 is None else d - t0
        if isinstance(key, np.timedelta64):
            key = key / np.timedelta64(1, "s")
        if np.any(ok):
            out[key] = (lo[ok], la[ok], v[ok], err[ok])
    return out
