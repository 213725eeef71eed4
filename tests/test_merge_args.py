"""The merge API rejects bad arguments before any device call (runs on the CPU):
a mask of the wrong shape, host or wrong-size or wrong-dtype buffers, and
blend inputs that are not float64 device tensors of one shape.  The host
triangle tables of irregular subwindows, evaluated in the device's order,
equal LinearNDInterpolator bit for bit (run_assimilation.py:436-446)."""
import numpy as np
import pytest
import torch

from vardyn_amd._lib import VardynError
from vardyn_amd.merge import SubwindowMerge, blend, delaunay_tables


def test_mask_shape_checked_before_create():
    lon, lat = np.meshgrid(np.arange(8.0), np.arange(6.0))
    subs = [(np.arange(5.0), np.arange(4.0))]
    with pytest.raises(VardynError, match="mask shape"):
        SubwindowMerge(lon, lat, subs, mask=np.zeros((6, 7), dtype=bool))


def fake_merger():
    m = object.__new__(SubwindowMerge)
    m.device = torch.device("cuda", 0)
    m.ny, m.nx, m.sub_points, m.nsub, m._h = 6, 8, 20, 1, None
    return m


def test_merge_packed_rejects_bad_buffers():
    m = fake_merger()
    host = torch.zeros(2 * 20, dtype=torch.float64)
    with pytest.raises(VardynError, match="contiguous float64 tensor on cuda"):
        m.merge_packed(host, 2, out=torch.zeros((2, 6, 8), dtype=torch.float64))
    with pytest.raises(VardynError, match="T must be"):
        m.merge_packed(host, 0)


def test_blend_rejects_host_and_mismatched_tensors():
    a = torch.zeros(10, dtype=torch.float64)
    with pytest.raises(VardynError, match="device tensor"):
        blend(a, a.clone(), 0.5, 0.5)
    with pytest.raises(VardynError, match="device tensor"):
        blend(a.float(), a.float(), 0.5, 0.5)


def test_irregular_subwindow_shape_checked():
    lon, lat = np.meshgrid(np.arange(8.0), np.arange(6.0))
    with pytest.raises(VardynError, match="irregular subwindow"):
        SubwindowMerge(lon, lat, [(np.arange(5.0), np.arange(4.0))], geo_grid=[False])
    with pytest.raises(VardynError, match="one flag per subwindow"):
        SubwindowMerge(lon, lat, [(np.arange(5.0), np.arange(4.0))], geo_grid=[True, False])


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_delaunay_tables_equal_linearnd_bitwise(seed):
    from oracle.merge import _interp_tri
    rng = np.random.default_rng(seed)
    i, j = np.meshgrid(np.arange(13 + seed), np.arange(11))
    lo = 0.5 + 0.4 * (i + 0.2 * j) + 0.05 * rng.standard_normal(i.shape)
    la = 40.5 + 0.4 * (j - 0.15 * i) + 0.05 * rng.standard_normal(i.shape)
    lon, lat = np.meshgrid(np.linspace(0.0, 10.0, 57), np.linspace(40.0, 48.0, 43))
    # targets on subwindow nodes too
    lon[0, :lo.shape[1]], lat[0, :lo.shape[1]] = lo[3], la[3]
    v = rng.standard_normal(lo.shape)
    ref = _interp_tri(lo, la, v, lat, lon).ravel()
    idx, w = delaunay_tables(lo, la, lon, lat)
    inside = idx[:, 0] >= 0
    assert np.array_equal(inside, ~np.isnan(ref)) and inside.sum() > 100
    acc = np.zeros(inside.sum())
    for k in range(3):   # the device's order: ((0 + c0 v0) + c1 v1) + c2 v2
        acc = acc + w[inside, k] * v.ravel()[idx[inside, k]]
    assert np.array_equal(acc, ref[inside])
