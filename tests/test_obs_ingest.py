"""Observation ingestion (vardyn_amd/obs.py, after mapping/src/obs.py:172-330
and the merge of obsop.py:513-547): hand-worked known answers for every step
-- longitude conventions, the box on a swath, the inclusive time windows (a
sample on a boundary goes to both dates), MDT interpolation (exact for a
linear field, also when extrapolating), noise, varmax, coarsening means with
trimming and NaN, error fields, and the per-date merge."""
import numpy as np
import pytest

from vardyn_amd.obs import _coarsen_mean, convert_lon, grid_interpolator, merge_dict_obs, obs_alti


def test_longitude_conventions():
    assert np.array_equal(convert_lon([-10.0, 5.0], "0_360"), [350.0, 5.0])
    assert np.array_equal(convert_lon([10.0, 200.0], "-180_180"), [10.0, -160.0])
    assert np.array_equal(convert_lon([-10.0, 5.0], "-180_180"), [-10.0, 5.0])
    assert np.array_equal(convert_lon([350.0, 5.0], "0_360"), [350.0, 5.0])


def test_time_windows_inclusive_and_half_open():
    t = np.array([0.0, 5.0, 10.0, 15.0, 20.0])
    v = np.arange(5.0)
    out = obs_alti(t, 1.0 + 0 * t, 2.0 + 0 * t, v, [5.0, 15.0, 40.0], 10.0)
    assert sorted(out) == [5.0, 15.0]
    assert np.array_equal(out[5.0][2], [0.0, 1.0, 2.0])      # [0, 10], both ends
    assert np.array_equal(out[15.0][2], [2.0, 3.0, 4.0])     # 10 is in both windows
    # unsorted times: label slicing is impossible, the half-open window is used
    tu = np.array([10.0, 0.0, 5.0, 20.0, 15.0])
    out = obs_alti(tu, 1.0 + 0 * tu, 2.0 + 0 * tu, v, [5.0, 15.0], 10.0)
    assert np.array_equal(out[5.0][2], [1.0, 2.0])           # [0, 10)
    assert np.array_equal(out[15.0][2], [0.0, 4.0])          # [10, 20)


def test_box_on_a_swath():
    t = np.array([0.0, 1.0, 2.0])
    lon = np.array([[0.0, 1.0, 2.0, 3.0], [0.5, 1.5, 2.5, 3.5], [9.0, 9.0, 9.0, 9.0]])
    lat = np.zeros_like(lon)
    v = np.arange(12.0).reshape(3, 4)
    out = obs_alti(t, lon, lat, v, [1.0], 10.0, bbox=[0.5, 2.6, -1.0, 1.0])
    lo, la, vv, err = out[1.0]
    # time row 2 has no pixel inside and is dropped; column 3 has none either
    assert vv.shape == (2, 3) and err is None
    want = np.array([[np.nan, 1.0, 2.0], [4.0, 5.0, 6.0]])
    assert np.array_equal(np.isnan(vv), np.isnan(want)) and np.allclose(vv[~np.isnan(want)], want[~np.isnan(want)])
    assert np.isnan(lo[0, 0]) and lo[1, 0] == 0.5


@pytest.mark.parametrize("layout", ["lat_lon", "lon_lat"])
def test_mdt_interpolation_exact_for_linear_fields(layout):
    glon, glat = np.linspace(-5, 5, 11), np.linspace(30, 40, 6)
    f = lambda lo, la: 0.1 + 0.02 * lo - 0.03 * la
    field = f(glon[None, :], glat[:, None])
    interp = grid_interpolator(glon, glat, field if layout == "lat_lon" else field.T)
    t = np.array([0.0, 1.0, 2.0])
    lon, lat = np.array([-4.3, 0.25, 7.0]), np.array([31.1, 35.0, 42.5])   # last one extrapolated
    y = np.array([1.0, 2.0, 3.0])
    add = obs_alti(t, lon, lat, y, [1.0], 10.0, mdt=interp, add_mdt=True)[1.0][2]
    sub = obs_alti(t, lon, lat, y, [1.0], 10.0, mdt=interp, substract_mdt=True)[1.0][2]
    assert np.allclose(add, y + f(lon, lat), atol=1e-13)
    assert np.allclose(sub, y - f(lon, lat), atol=1e-13)


def test_noise_varmax_and_combination():
    t = np.zeros(20000)
    out = obs_alti(t, 0 * t, 0 * t, 0 * t, [0.0], 1.0, synthetic_noise=0.05,
                   rng=np.random.default_rng(3))[0.0][2]
    assert abs(out.mean()) < 3e-3 and abs(out.std() - 0.05) < 2e-3
    again = obs_alti(t, 0 * t, 0 * t, 0 * t, [0.0], 1.0, synthetic_noise=0.05,
                     rng=np.random.default_rng(3))[0.0][2]
    assert np.array_equal(out, again)
    t = np.zeros(4)
    v = obs_alti(t, 0 * t, 0 * t, np.array([0.1, -2.0, 0.5, 3.0]), [0.0], 1.0, varmax=1.0)[0.0][2]
    assert np.array_equal(np.isnan(v), [False, True, False, True])
    a, b = np.array([1.0, 2.0, 3.0, 4.0]), np.array([0.5, 0.5, 1.0, 1.0])
    v = obs_alti(t, 0 * t, 0 * t, [a, b], [0.0], 1.0, combine_signs=[1, -1])[0.0][2]
    assert np.array_equal(v, a - b)


def test_coarsening_means():
    assert np.array_equal(_coarsen_mean([1.0, 2.0, 3.0, 4.0, 5.0], 2), [1.5, 3.5])
    got = _coarsen_mean([1.0, np.nan, 3.0, 4.0, np.nan, np.nan], 2)
    assert got[0] == 1.0 and got[1] == 3.5 and np.isnan(got[2])
    a = np.arange(24.0).reshape(4, 6)
    want = np.array([[(0 + 1 + 6 + 7) / 4, (2 + 3 + 8 + 9) / 4, (4 + 5 + 10 + 11) / 4],
                     [(12 + 13 + 18 + 19) / 4, (14 + 15 + 20 + 21) / 4, (16 + 17 + 22 + 23) / 4]])
    assert np.array_equal(_coarsen_mean(a, 2), want)
    # subsampling applies to values and positions alike
    t = np.arange(6.0)
    lo, la, v, _ = obs_alti(t, t * 0.1, t * 0.2, t * 10.0, [2.5], 10.0, subsampling=3)[2.5]
    assert np.allclose(lo, [0.1, 0.4]) and np.allclose(la, [0.2, 0.8]) and np.allclose(v, [10.0, 40.0])


def test_error_field_and_merge():
    glon, glat = np.linspace(-5, 5, 11), np.linspace(30, 40, 6)
    err = grid_interpolator(glon, glat, 0.02 + 0 * glon[None, :] + 0 * glat[:, None])
    t = np.array([0.0, 1.0])
    d1 = obs_alti(t, np.array([0.0, 1.0]), np.array([35.0, 36.0]), np.array([1.0, np.nan]), [0.5], 2.0,
                  err_field=err, facR=2.0)
    d2 = obs_alti(t, np.array([2.0, 3.0]), np.array([35.0, 36.0]), np.array([3.0, 4.0]), [0.5, 5.0], 2.0)
    merged = merge_dict_obs([d1, d2], sigma_noise=[None, 0.01], t0=-1.5)
    assert list(merged) == [2.0]                           # 0.5 - (-1.5); no data at 5.0
    lo, la, y, e = merged[2.0]
    assert np.array_equal(y, [1.0, 3.0, 4.0])              # NaN dropped, dataset order kept
    assert np.array_equal(lo, [0.0, 2.0, 3.0])
    assert np.allclose(e, [0.04, 0.01, 0.01])              # field x facR, then sigma_noise


def test_round_trip_of_a_synthetic_observation_set():
    """Raw along-track samples of a synthetic observation set (time, lon, lat,
    value, error per sample), ingested on the model's steps and merged, give
    back the observation dictionary the problem builder makes from it."""
    from vardyn_amd import synthetic as syn
    case = syn.QgCase("t", 40, 48, 10e3, 2.7, 1200.0, 60, n_tracks=5, track_km=7.0, seed=3)
    grid = syn.make_grid(case.ny, case.nx, case.dx)
    obs = syn.make_observations(case, grid)
    T = np.arange(case.n_steps + 1) * case.dt
    rng = np.random.default_rng(8)
    for e in obs.steps.values():   # values of a truth run are not needed here: seeded ones
        e["y"] = rng.standard_normal(e["lon"].size)
    want = {float(T[s]): (e["lon"], e["lat"], e["y"], e["err"]) for s, e in obs.steps.items()}
    # one raw nadir dataset: every sample with its own time stamp
    tt = np.concatenate([np.full(v[0].size, t) for t, v in want.items()])
    lon, lat, y, err = (np.concatenate([v[k] for v in want.values()]) for k in range(4))
    from scipy.interpolate import NearestNDInterpolator
    errf = NearestNDInterpolator(np.column_stack([lon, lat]), err)
    got = merge_dict_obs([obs_alti(tt, lon, lat, y, T, case.dt, err_field=lambda p: errf(*p))])
    assert list(got) == list(want)
    for t in want:
        for a, b in zip(got[t], want[t]):
            assert np.array_equal(a, b)


def test_read_netcdf_classic_files(tmp_path):
    """read_obs_files (the file selection and CF decoding in front of
    _obs_alti, obs.py:120-147): two netCDF classic files of one nadir track
    written out of time order, SSH packed as int16 with scale_factor,
    add_offset and _FillValue, times as "hours since"; the decoded dataset
    equals the arrays it was written from (fill values NaN, files ordered by
    their first time), and obs_alti gives the same result from it as from
    those arrays."""
    from scipy.io import netcdf_file
    from vardyn_amd.obs import read_obs_files
    rng = np.random.default_rng(4)
    n = 40
    hours = np.arange(2 * n, dtype=np.float64) * 0.25
    lon = -60 + 0.05 * np.arange(2 * n)
    lat = 30 + 0.04 * np.arange(2 * n)
    packed = rng.integers(-2000, 2000, 2 * n).astype(np.int16)
    packed[[3, 47]] = -32767
    scale, offset = 0.001, 0.5
    for k, sl in enumerate((slice(n, 2 * n), slice(0, n))):  # the later half is file 0
        with netcdf_file(str(tmp_path / f"alti_{k}.nc"), "w") as nc:
            nc.createDimension("time", n)
            t = nc.createVariable("time", "f8", ("time",))
            t[:] = hours[sl]
            t.units = b"hours since 2012-10-01 00:00:00"
            for name, vals in (("longitude", lon[sl]), ("latitude", lat[sl])):
                v = nc.createVariable(name, "f8", ("time",))
                v[:] = vals
            v = nc.createVariable("sla", "i2", ("time",))
            v[:] = packed[sl]
            v.scale_factor = scale
            v.add_offset = offset
            v._FillValue = np.int16(-32767)
    ds = read_obs_files(str(tmp_path / "alti_"), "time", "longitude", "latitude", ["sla"])
    t_ref = np.datetime64("2012-10-01T00:00:00", "ns") + (hours * 3600e9).astype(np.int64).astype("timedelta64[ns]")
    np.testing.assert_array_equal(ds["time"], t_ref)
    np.testing.assert_array_equal(ds["longitude"], lon)
    # scipy stores Python float attributes as float32; they are applied in FP64
    sla = packed.astype(np.float64) * float(np.float32(scale)) + float(np.float32(offset))
    sla[[3, 47]] = np.nan
    np.testing.assert_array_equal(ds["sla"], sla)
    dates = np.datetime64("2012-10-01T00:00", "ns") + np.arange(0, 20, 3) * np.timedelta64(3600, "s")
    step = np.timedelta64(3 * 3600, "s")
    a = obs_alti(ds["time"], ds["longitude"], ds["latitude"], ds["sla"], dates, step)
    b = obs_alti(t_ref, lon, lat, sla, dates, step)
    assert a.keys() == b.keys() and len(a) > 3
    for d in a:
        for x, y in zip(a[d][:3], b[d][:3]):
            np.testing.assert_array_equal(x, y)
    with pytest.raises(FileNotFoundError):
        read_obs_files(str(tmp_path / "missing_"), "time", "longitude", "latitude", ["sla"])


def test_read_auxdata_classic_file(tmp_path):
    """read_auxdata (tools.py:204-234) from a netCDF classic file: a linear
    MDT on (lat, lon) axes with longitudes in 0..360 is reproduced exactly at
    any point by the interpolator it returns, under the -180_180 convention."""
    from scipy.io import netcdf_file
    from vardyn_amd.obs import read_auxdata
    lon = np.arange(290.0, 310.0, 0.5)
    lat = np.arange(25.0, 40.0, 0.5)
    mdt = 0.01 * (lon[None, :] - 300.0) + 0.02 * lat[:, None]
    f = str(tmp_path / "mdt.nc")
    with netcdf_file(f, "w") as nc:
        nc.createDimension("lat", lat.size)
        nc.createDimension("lon", lon.size)
        nc.createVariable("lon", "f8", ("lon",))[:] = lon
        nc.createVariable("lat", "f8", ("lat",))[:] = lat
        nc.createVariable("mdt", "f8", ("lat", "lon"))[:] = mdt
    fi = read_auxdata(f, {"lon": "lon", "lat": "lat", "mdt": "mdt"}, "-180_180")
    pts = np.array([[-65.2, 31.3], [-50.75, 26.0], [-69.0, 39.2]])
    np.testing.assert_allclose(fi(pts), 0.01 * (pts[:, 0] + 360.0 - 300.0) + 0.02 * pts[:, 1], rtol=1e-12)
