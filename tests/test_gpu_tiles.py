"""QgmWithTiles on the GPU (per-tile plans, vd_tiles_* gather/blend kernels)
against the FP64 tiled oracle (jqgm.py:1033-1258).

Bars: FP64 step, tangent-linear and adjoint within 1e-10 relative of the oracle
(the sine transforms are GEMMs here and FFTs in the oracle), with the tiles in
shared launches (ragged group) and tile by tile; the two engines within 1e-12
of each other on the reference's default tiling (32-point tiles, overlap 16); the gather/blend
kernels exactly adjoint to each other up to rounding (1e-13); blending the tiles
of a field gives the field back (1e-14).
"""
import numpy as np
import pytest
import torch

from oracle.tiles import TiledQgmOracle
from vardyn_amd import _lib
from vardyn_amd.qg import current_stream_ptr, ptr
from vardyn_amd.tiles import QgmWithTiles

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    assert np.array_equal(np.isnan(a), np.isnan(b)), "NaN pattern differs"
    m = ~np.isnan(b)
    return np.linalg.norm(a[m] - b[m]) / max(np.linalg.norm(b[m]), 1e-300)


def pair(ny, nx, scheme, ts, ov, land=False, mdt=False, kdiff=False, fvar=False, seed=0,
         engine="group"):
    rng = np.random.default_rng(seed)
    dx = 10e3
    ssh = np.zeros((ny, nx))
    if land:
        ssh[ny // 2, nx // 3] = np.nan
        ssh[ny // 3, nx // 2:nx // 2 + 2] = np.nan
    f = 1e-4 + (2e-6 * rng.standard_normal((ny, nx)) if fvar else 0.0)
    c = 2.0 + 0.5 * rng.random((ny, nx))
    m = 0.02 * rng.standard_normal((ny, nx)) if mdt else None
    K = 30.0 if kdiff else None
    orc = TiledQgmOracle(dx, dx, 1200.0, (ny, nx), c=c, f=f, SSH=ssh, Kdiffus=K, mdt=m,
                         time_scheme=scheme, tile_size=ts, tile_overlap=ov)
    gpu = QgmWithTiles(dx=dx * np.ones((ny, nx)), dy=dx * np.ones((ny, nx)), dt=1200.0, SSH=ssh,
                       c=c, Kdiffus=K, f=f, time_scheme=scheme, mdt=m, tile_size=ts,
                       tile_overlap=ov, engine=engine)
    h0 = 0.05 * rng.standard_normal((ny, nx))
    h0[np.isnan(ssh)] = np.nan
    hb = 0.01 * rng.standard_normal((ny, nx))
    return orc, gpu, h0, hb, rng, np.isnan(ssh)


@pytest.mark.parametrize("engine", ["group", "plans"])
@pytest.mark.parametrize("scheme", ["Euler", "rk2"])
@pytest.mark.parametrize("opts", [dict(), dict(land=True, mdt=True, kdiff=True, fvar=True)])
def test_step_tgl_adj_against_oracle(scheme, opts, engine):
    orc, gpu, h0, hb, rng, land = pair(76, 58, scheme, 32, 8, engine=engine, **opts)
    n = 6
    h1 = gpu.step(h0, hb, n)
    assert isinstance(h1, np.ndarray)
    assert rel(h1, orc.step(h0, hb, n)) < 1e-10
    dh = 1e-3 * rng.standard_normal(h0.shape)
    dh[land] = 0
    assert rel(gpu.step_tgl(dh, h0, hb, n), orc.step_tgl(dh, h0, hb, n)) < 1e-10
    ad = rng.standard_normal(h0.shape)
    ad[land] = 0
    assert rel(gpu.step_adj(ad, h0, hb, n), orc.step_adj(ad, h0, hb, n)) < 1e-10


def test_dropin_attributes_and_torch_io():
    orc, gpu, h0, hb, _, _ = pair(40, 40, "Euler", 24, 6)
    assert gpu.step_jit == gpu.step and gpu.step_adj_jit == gpu.step_adj
    h1 = gpu.step_jit(torch.as_tensor(h0, device="cuda"), torch.as_tensor(hb, device="cuda"), 3)
    assert isinstance(h1, torch.Tensor) and h1.is_cuda
    assert rel(h1.cpu().numpy(), orc.step(h0, hb, 3)) < 1e-10


def test_gather_blend_kernels():
    _, gpu, _, _, rng, _ = pair(70, 90, "Euler", 30, 8)
    x = torch.as_tensor(rng.standard_normal((70, 90)), device="cuda")
    y = torch.as_tensor(rng.standard_normal((70, 90)), device="cuda")
    F = torch.as_tensor(rng.standard_normal((gpu.ntile, 30, 30)), device="cuda")
    # blending the tiles of a field returns the field
    assert rel(gpu._blend(gpu._gather(x)).cpu().numpy(), x.cpu().numpy()) < 1e-14
    s = current_stream_ptr(gpu.device)
    # <gather x, F> = <x, gather^T F>
    gT = torch.empty_like(x)
    _lib.call("vd_tiles_gather_adj", *gpu._dims(), ptr(F), ptr(gT), s)
    lhs, rhs = float((gpu._gather(x) * F).sum()), float((x * gT).sum())
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    # <blend F, y> = <F, blend^T y>
    bT = torch.empty_like(F)
    _lib.call("vd_tiles_blend_adj", *gpu._dims(), ptr(gpu._w), ptr(gpu._wsum), ptr(y), ptr(bT), s)
    lhs, rhs = float((gpu._blend(F) * y).sum()), float((F * bT).sum())
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


@pytest.mark.parametrize("scheme", ["Euler", "rk2"])
def test_group_engine_matches_tile_by_tile(scheme):
    """The reference's default tiling (tile_size 32, tile_overlap 16; config_default.py:398-400):
    shared launches against tile by tile, with land, over several steps."""
    _, grp, h0, hb, rng, land = pair(96, 80, scheme, 32, 16, land=True)
    _, one = pair(96, 80, scheme, 32, 16, land=True, engine="plans")[:2]
    assert grp.ntile == one.ntile == 6 * 5
    n = 5
    assert rel(grp.step(h0, hb, n), one.step(h0, hb, n)) < 1e-12
    ad = rng.standard_normal(h0.shape)
    ad[land] = 0
    assert rel(grp.step_adj(ad, h0, hb, n), one.step_adj(ad, h0, hb, n)) < 1e-12
