"""GPU subwindow / time-window merge against the oracle
(run_assimilation.py:351-446, 537-607, 732-830): weights, weight sums and
merged fields within 1e-13 relative, with the oracle's NaN pattern, for
regular and irregular (Delaunay) subwindow grids; the blend and its
weights."""
import numpy as np
import pytest
import torch

from oracle.merge import merge_date, time_blend, weights_map

pytestmark = pytest.mark.gpu
TOL = 1e-13


def _case(seed, nsub, ny=37, nx=53, irregular=True):
    rng = np.random.default_rng(seed)
    lon, lat = np.meshgrid(np.linspace(-4.0, 6.0, nx), np.linspace(30.0, 38.0, ny))
    subs = []
    for s in range(nsub):
        lo0, la0 = rng.uniform(-5.0, 3.0), rng.uniform(29.0, 35.0)
        mx, my = int(rng.integers(6, 30)), int(rng.integers(5, 25))
        if irregular:
            lo1 = lo0 + np.cumsum(rng.uniform(0.1, 0.5, mx))
            la1 = la0 + np.cumsum(rng.uniform(0.1, 0.5, my))
        else:
            lo1 = np.linspace(lo0, lo0 + 5.0, mx)
            la1 = np.linspace(la0, la0 + 4.0, my)
        subs.append((lo1, la1))
    return lon, lat, subs, rng


def _close(a, b, tol=TOL):
    a, b = np.asarray(a), np.asarray(b)
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb)
    ok = ~na
    scale = max(np.abs(b[ok]).max(), 1e-300) if ok.any() else 1.0
    assert np.abs(a[ok] - b[ok]).max() / scale <= tol if ok.any() else True


@pytest.mark.parametrize("seed,nsub,T,irregular,masked,nan_fields", [
    (1, 1, 3, False, False, False),
    (2, 4, 11, True, False, False),
    (3, 6, 8, True, True, True),
    (4, 9, 17, False, True, True),
])
def test_merge_matches_oracle(seed, nsub, T, irregular, masked, nan_fields):
    from vardyn_amd.merge import SubwindowMerge
    lon, lat, subs, rng = _case(seed, nsub, irregular=irregular)
    mask = (rng.random(lon.shape) < 0.05) if masked else None
    w_ref, ws_ref = weights_map(lon, lat, subs)
    m = SubwindowMerge(lon, lat, subs, mask=mask)
    _close(m.weights_space.cpu().numpy(), np.stack(w_ref))
    _close(m.weights_space_sum.cpu().numpy(), ws_ref)
    fields = []
    for lo1, la1 in subs:
        f = rng.standard_normal((T, la1.size, lo1.size))
        if nan_fields:
            f[rng.random(f.shape) < 0.03] = np.nan
        fields.append(f)
    out = m.merge([torch.from_numpy(f).cuda() for f in fields]).cpu().numpy()
    for t in range(T):
        ref = merge_date(lon, lat, subs, [f[t] for f in fields], w_ref, ws_ref, mask=mask)
        _close(out[t], ref)


def _curvilinear(rng, lon0, lat0, nx, ny, d=0.35):
    """An irregular (GRID_CAR) subwindow grid: a sheared, jittered lattice."""
    i, j = np.meshgrid(np.arange(nx), np.arange(ny))
    lo = lon0 + d * (i + 0.2 * j) + 0.04 * rng.standard_normal(i.shape)
    la = lat0 + d * (j - 0.15 * i) + 0.04 * rng.standard_normal(i.shape)
    return lo, la


@pytest.mark.parametrize("seed,kinds,T,masked,nan_fields", [
    (11, "c", 4, False, False),
    (12, "cgc", 9, True, True),
    (13, "ccgcg", 13, False, True),
])
def test_irregular_subwindows_match_oracle(seed, kinds, T, masked, nan_fields):
    """Irregular subwindows (run_assimilation.py:436-446: LinearNDInterpolator
    over the subwindow's points), alone and mixed with regular ones."""
    from vardyn_amd.merge import SubwindowMerge
    rng = np.random.default_rng(seed)
    lon, lat = np.meshgrid(np.linspace(-4.0, 6.0, 61), np.linspace(30.0, 38.0, 47))
    subs, geo = [], []
    for k in kinds:
        lo0, la0 = rng.uniform(-5.0, 2.0), rng.uniform(29.0, 34.0)
        if k == "c":
            subs.append(_curvilinear(rng, lo0, la0, int(rng.integers(8, 20)), int(rng.integers(7, 16))))
        else:
            subs.append((np.linspace(lo0, lo0 + 5.0, int(rng.integers(6, 25))),
                         np.linspace(la0, la0 + 4.0, int(rng.integers(5, 20)))))
        geo.append(k == "g")
    mask = (rng.random(lon.shape) < 0.05) if masked else None
    w_ref, ws_ref = weights_map(lon, lat, subs, geo=geo)
    assert sum(np.isfinite(w).sum() for w in w_ref) > 200
    m = SubwindowMerge(lon, lat, subs, mask=mask, geo_grid=geo)
    _close(m.weights_space.cpu().numpy(), np.stack(w_ref))
    _close(m.weights_space_sum.cpu().numpy(), ws_ref)
    fields = []
    for (lo, la), g in zip(subs, geo):
        shape = (la.size, lo.size) if g else lo.shape
        f = rng.standard_normal((T,) + shape)
        if nan_fields:
            f[rng.random(f.shape) < 0.03] = np.nan
        fields.append(f)
    out = m.merge([torch.from_numpy(f).cuda() for f in fields]).cpu().numpy()
    for t in range(T):
        ref = merge_date(lon, lat, subs, [f[t] for f in fields], w_ref, ws_ref, mask=mask, geo=geo)
        _close(out[t], ref)


def test_target_on_subwindow_nodes():
    """Targets exactly on nodes and on the last node of an axis (scipy's closed last cell)."""
    from vardyn_amd.merge import SubwindowMerge
    lon, lat = np.meshgrid(np.linspace(0.0, 4.0, 9), np.linspace(10.0, 12.0, 5))
    subs = [(np.linspace(0.0, 4.0, 5), np.linspace(10.0, 12.0, 3)),
            (np.linspace(2.0, 4.0, 3), np.linspace(10.0, 12.0, 5))]
    rng = np.random.default_rng(9)
    fields = [rng.standard_normal((2, la.size, lo.size)) for lo, la in subs]
    fields[0][:, 1, 2] = np.nan
    w_ref, ws_ref = weights_map(lon, lat, subs)
    m = SubwindowMerge(lon, lat, subs)
    out = m.merge([torch.from_numpy(f).cuda() for f in fields]).cpu().numpy()
    for t in range(2):
        _close(out[t], merge_date(lon, lat, subs, [f[t] for f in fields], w_ref, ws_ref))


def test_repeatable_and_packed():
    from vardyn_amd.merge import SubwindowMerge
    lon, lat, subs, rng = _case(5, 5)
    m = SubwindowMerge(lon, lat, subs)
    fields = [torch.from_numpy(rng.standard_normal((9, la.size, lo.size))).cuda() for lo, la in subs]
    a = m.merge(fields)
    packed, T = m.pack(fields)
    b = m.merge_packed(packed, T)
    assert torch.equal(a.nan_to_num(7.0), b.nan_to_num(7.0))


def test_blend_and_time_weights():
    from vardyn_amd.merge import blend, time_window_weights
    rng = np.random.default_rng(2)
    a = torch.from_numpy(rng.standard_normal((3, 20, 30))).cuda()
    b = torch.from_numpy(rng.standard_normal((3, 20, 30))).cuda()
    starts, ends = [0.0, 10.0, 30.0], [20.0, 40.0, 50.0]
    for date in (5.0, 10.0, 12.5, 15.0, 20.0, 25.0, 35.0, 50.0, 60.0):
        got = time_window_weights(date, starts, ends)
        active = [i for i in range(3) if starts[i] <= date <= ends[i]]
        if not active:
            assert got == []
            continue
        if len(active) == 1:
            assert got == [(active[0], 1.0)]
            continue
        i, j = active[:2]
        w1, w2 = time_blend(date, starts[j], ends[i])
        assert got[0][0] == i and got[1][0] == j
        assert abs(got[0][1] - w1) < 1e-15 and abs(got[1][1] - w2) < 1e-15
        out = blend(a, b, got[0][1], got[1][1]).cpu().numpy()
        ref = w1 * a.cpu().numpy() + w2 * b.cpu().numpy()
        assert np.array_equal(out, ref)


def test_rejects_bad_axes():
    from vardyn_amd._lib import VardynError
    from vardyn_amd.merge import SubwindowMerge
    lon, lat = np.meshgrid(np.linspace(0.0, 1.0, 4), np.linspace(0.0, 1.0, 3))
    with pytest.raises(VardynError):
        SubwindowMerge(lon, lat, [(np.array([0.0, 0.5, 0.4]), np.array([0.0, 1.0]))])
    with pytest.raises(VardynError):
        SubwindowMerge(lon, lat, [(np.array([0.0]), np.array([0.0, 1.0]))])
