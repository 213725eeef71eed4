"""Host logic of the shallow-water wrapper without a GPU: parameters the GPU
model does not implement are rejected before any device call, and the C ABI
validates its parameters before touching a device."""
import ctypes as C

import numpy as np
import pytest

from vardyn_amd import _lib
from vardyn_amd._lib import SwParams, VardynError
from vardyn_amd.sw import SW


def param(**kw):
    nx, ny, nl = 12, 10, 2
    p = dict(nx=nx, ny=ny, nl=nl, dx=8e3, dy=8e3, H=np.array([500., 1000.]).reshape(2, 1, 1),
             g_prime=np.array([0.02, 0.01]), f=np.full((nx + 1, ny + 1), 1e-4), dt=120.0)
    p.update(kw)
    return p


@pytest.mark.parametrize("kw", [dict(taux=0.1),
                                dict(taux=0.1, h_wind=np.ones((3, 3))), dict(tauy=np.ones((12, 10)), h_wind=50.0),
                                dict(barotropic_filter=True), dict(n_ens=2),
                                dict(dx=np.linspace(8e3, 9e3, 12)), dict(H=np.ones((2, 12, 10))),
                                dict(f=np.zeros((12, 10))), dict(mask=np.zeros((12, 10))),
                                dict(flag_obc=True), dict(forcing_momentum="mass_consistent")])
def test_unsupported_parameters_are_rejected(kw):
    """Rejected before any device work: options not built, and wind stress
    without h_wind (the reference's fallback cannot broadcast with per-layer
    constant depths) or with mis-shaped tau / h_wind."""
    with pytest.raises(VardynError):
        SW(param(**kw))


def test_capi_validates_without_a_device():
    lib = _lib.load()
    H = np.array([500.0]); gp = np.array([0.02]); f = np.full((13, 11), 1e-4)
    p = SwParams(12, 10, 1, 8e3, 8e3, 120.0, H.ctypes.data, gp.ctypes.data, f.ctypes.data, None,
                 0.0, 0.0, 0.1, 0.0, 1.0, 0.0)
    h = C.c_void_p()
    assert lib.vd_sw_create(C.byref(p), C.byref(h)) == -1
    assert "h_min_sharpness" in _lib.last_error()
    p.h_min_sharpness = 10.0
    p.nl = 9
    assert lib.vd_sw_create(C.byref(p), C.byref(h)) == -1
