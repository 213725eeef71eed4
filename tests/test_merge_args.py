"""The merge API rejects bad arguments before any device call (runs on the CPU):
a mask of the wrong shape, host or wrong-size or wrong-dtype buffers, and
blend inputs that are not float64 device tensors of one shape."""
import numpy as np
import pytest
import torch

from vardyn_amd._lib import VardynError
from vardyn_amd.merge import SubwindowMerge, blend


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
