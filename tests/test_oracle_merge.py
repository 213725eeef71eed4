"""Pins the oracle of the subwindow / time-window merge
(run_assimilation.py:351-446, 537-607, 732-830) by closed forms, exact
reproductions and brute force."""
import numpy as np

from oracle.merge import hann, merge_date, time_blend, weights_map


def _grid(lon0, lon1, nx, lat0, lat1, ny):
    return np.linspace(lon0, lon1, nx), np.linspace(lat0, lat1, ny)


def _target(nx=23, ny=19):
    lon, lat = np.meshgrid(np.linspace(0.0, 10.0, nx), np.linspace(40.0, 48.0, ny))
    return lon, lat


def test_hann_closed_form():
    w = hann(9)
    assert w[0] == 0 and abs(w[-1]) < 1e-15 and abs(w[4] - 1) < 1e-15
    assert np.allclose(w, np.sin(np.pi * np.arange(9) / 8) ** 2, atol=1e-15)


def test_linear_field_reproduced():
    lon, lat = _target()
    subs = [_grid(-1.0, 6.3, 17, 39.0, 45.1, 13), _grid(4.0, 11.0, 29, 43.0, 49.0, 11)]
    lin = lambda lo, la: 0.3 + 2.0 * lo - 0.7 * la
    fields = [lin(*np.meshgrid(lo1, la1)) for lo1, la1 in subs]
    w, ws = weights_map(lon, lat, subs)
    out = merge_date(lon, lat, subs, fields, w, ws)
    ok = ws > 0
    assert ok.sum() > 100
    assert np.allclose(out[ok], lin(lon, lat)[ok], rtol=0, atol=1e-12)


def test_partition_of_unity():
    lon, lat = _target()
    subs = [_grid(-1.0, 6.3, 17, 39.0, 45.1, 13), _grid(4.0, 11.0, 29, 43.0, 49.0, 11),
            _grid(2.0, 8.0, 7, 38.0, 50.0, 31)]
    w, ws = weights_map(lon, lat, subs)
    out = merge_date(lon, lat, subs, [np.full((la.size, lo.size), 3.25) for lo, la in subs], w, ws)
    ok = ws > 0
    assert np.allclose(out[ok], 3.25, rtol=1e-15, atol=0)
    # the weight sum equals the sum of the weights where they are defined
    tot = np.nansum(np.stack(w), axis=0)
    assert np.array_equal(ws, tot)


def test_single_subwindow_identity():
    lon, lat = _target()
    sub = [(lon[0], lat[:, 0])]
    w, ws = weights_map(lon, lat, sub)
    assert np.all(w[0] == 1) and np.all(ws == 1)
    f = np.random.default_rng(3).standard_normal(lon.shape)
    out = merge_date(lon, lat, sub, [f], w, ws, mask=None)
    assert np.allclose(out, f, rtol=0, atol=1e-14)
    m = np.zeros(lon.shape, bool)
    m[2, 3] = True
    out = merge_date(lon, lat, sub, [f], w, ws, mask=m)
    assert np.isnan(out[2, 3]) and np.isfinite(out[~m]).all()


def test_nan_corner_reaches_its_cells():
    lon, lat = _target(31, 27)
    lo1, la1 = _grid(-0.5, 10.5, 12, 39.5, 48.5, 10)
    f = np.ones((la1.size, lo1.size))
    f[4, 6] = np.nan
    w, ws = weights_map(lon, lat, [(lo1, la1)])
    out = merge_date(lon, lat, [(lo1, la1)], [f], w, ws)
    # brute force: cell i with la1[i] <= y < la1[i+1] (last cell closed)
    def cell(a, v):
        for i in range(a.size - 1):
            if a[i] <= v < a[i + 1] or (i == a.size - 2 and v == a[-1]):
                return i
        return -1
    for p in range(lon.size):
        i, j = cell(la1, lat.flat[p]), cell(lo1, lon.flat[p])
        touches = i in (3, 4) and j in (5, 6)
        assert (out.flat[p] == 0) == touches, p  # NaN values are skipped: nothing added
        if not touches:
            assert abs(out.flat[p] - 1) < 1e-14


def test_time_blend_weights():
    assert time_blend(10.0, 10.0, 20.0) == (1.0, 0.0)
    w1, w2 = time_blend(20.0, 10.0, 20.0)
    assert abs(w1) < 1e-15 and abs(w2 - 1) < 1e-15
    w1, w2 = time_blend(15.0, 10.0, 20.0)
    assert abs(w1 - 0.5) < 1e-15 and abs(w2 - 0.5) < 1e-15
    assert time_blend(5.0, 5.0, 5.0) == (0.5, 0.5)
    w1, w2 = time_blend(12.5, 10.0, 20.0)
    assert abs(w2 - np.sin(np.pi * 0.25 / 2) ** 2) < 1e-15 and abs(w1 + w2 - 1) < 1e-15


def _irregular(seed, lon0, lat0, nx, ny, d=0.4):
    """A curvilinear subwindow grid: a rotated, sheared and jittered lattice."""
    rng = np.random.default_rng(seed)
    i, j = np.meshgrid(np.arange(nx), np.arange(ny))
    lo = lon0 + d * (i + 0.2 * j) + 0.05 * rng.standard_normal(i.shape)
    la = lat0 + d * (j - 0.15 * i) + 0.05 * rng.standard_normal(i.shape)
    return lo, la


def test_irregular_branch_linear_exact_and_hull():
    from scipy.spatial import ConvexHull
    lon, lat = _target()
    subs = [_irregular(1, 0.5, 40.5, 14, 12), _irregular(2, 4.0, 43.0, 11, 10)]
    geo = [False, False]
    lin = lambda lo, la: 0.3 + 2.0 * lo - 0.7 * la
    w, ws = weights_map(lon, lat, subs, geo=geo)
    out = merge_date(lon, lat, subs, [lin(lo, la) for lo, la in subs], w, ws, geo=geo)
    ok = ws > 0
    assert ok.sum() > 50
    assert np.allclose(out[ok], lin(lon, lat)[ok], rtol=0, atol=1e-12)
    for (lo, la), wi in zip(subs, w):
        hull = ConvexHull(np.column_stack([lo.ravel(), la.ravel()]))
        pts = np.column_stack([lon.ravel(), lat.ravel()])
        dist = pts @ hull.equations[:, :2].T + hull.equations[:, 2]
        inside = np.all(dist <= 1e-12, axis=1).reshape(lon.shape)
        outside = np.any(dist > 1e-9, axis=1).reshape(lon.shape)
        assert np.all(~np.isnan(wi[inside])) and np.all(np.isnan(wi[outside]))
