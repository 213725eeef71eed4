"""The GPU 3-D wavelet basis (vardyn_amd/basis_wavelet3d.py, vd_w3_* in
vd_w3.hip) against the FP64 oracle (oracle/basis_wavelet3d.py).

Bars: the host tables (centres, time blocks, counts, Q) equal the oracle's;
operg within 1e-12 relative of the oracle's CSC synthesis at every basis time
(one MFMA product per tile, summed in a different order than scipy's CSC
product); operg_transpose within 1e-12; the two GPU operators adjoint to
each other within 1e-12; repeated calls bitwise equal; batches of more than
16 times (two passes) equal the per-time calls.  Cases: grids whose sides are
not multiples of the 16-point tile, a masked grid, and the 256 x 256 grid of
the 3-layer shallow-water configuration (4 scales, 48k coefficients).
"""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle.basis_wavelet3d import KM2DEG, BasisWavelet3dOracle
from vardyn_amd.basis_wavelet3d import Basis_wavelet3d

pytestmark = pytest.mark.gpu


def grid(ny, nx, lat0=40.0, dkm=8.0):
    lat1 = lat0 + (np.arange(ny) - (ny - 1) / 2) * dkm * KM2DEG
    lon1 = 3.0 + (np.arange(nx) - (nx - 1) / 2) * dkm * KM2DEG / np.cos(np.radians(lat0))
    return np.meshgrid(lon1, lat1)


def pair(ny, nx, mask=None, times=np.arange(0, 9.0), **kw):
    kw = {**dict(lmin=40.0, lmax=170.0, npsp=3.5, facpsp=3.5, tmin=4.0, tmax=4.0), **kw}
    lon, lat = grid(ny, nx)
    orc = BasisWavelet3dOracle(lon, lat, mask=mask, **kw)
    orc.set_basis(times)
    gpu = Basis_wavelet3d(SimpleNamespace(lon=lon, lat=lat, mask=mask), **kw)
    Xb, Q = gpu.set_basis(times, return_q=True)
    return orc, gpu, Q


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


CASES = {"ragged_tiles": dict(ny=37, nx=45), "masked": dict(ny=40, nx=33, masked=True),
         "small": dict(ny=16, nx=16)}


@pytest.mark.parametrize("case", list(CASES))
def test_tables_and_operators_match_oracle(case):
    c = dict(CASES[case])
    rng = np.random.default_rng(1)
    mask = (rng.random((c["ny"], c["nx"])) < 0.15) if c.pop("masked", False) else None
    orc, gpu, Q = pair(mask=mask, **c)
    assert gpu.nbasis == orc.nbasis and np.array_equal(gpu.NS, orc.NS) and gpu.ntheta == orc.ntheta
    assert np.all(Q == 0.1) and Q.size == orc.nbasis
    for f in range(orc.nfs):
        assert np.array_equal(gpu.ENSLON[f], orc.ENSLON[f]) and np.array_equal(gpu.ENSLAT[f], orc.ENSLAT[f])
    assert gpu.nnz_space == sum(g.nnz for g in orc.Gx)
    X = rng.standard_normal(orc.nbasis)
    times = orc.time
    fields = gpu.operg_batch(times, X).cpu().numpy()
    for it in range(times.size):
        ref = orc.operg(it, X)
        assert rel(fields[it], ref) < 1e-12, it
        if mask is not None:
            assert np.all(fields[it][mask] == 0.0)
    ad = rng.standard_normal((times.size,) + orc.shape_phys)
    got = gpu.operg_transpose_batch(times, ad).cpu().numpy()
    ref = sum(orc.operg_transpose(it, ad[it]) for it in range(times.size))
    assert rel(got, ref) < 1e-12
    # per-time calls with the reference's interface
    st = {"ssh": np.zeros(orc.shape_phys)}
    gpu.operg(times[3], X, State=st)
    assert rel(st["ssh"].cpu().numpy(), orc.operg(3, X)) < 1e-12
    adst = {"ssh": ad[5].copy()}
    g5 = gpu.operg_transpose(times[5], adst)
    assert rel(g5, orc.operg_transpose(5, ad[5])) < 1e-12 and np.all(adst["ssh"] == 0)


def test_adjointness_determinism_and_passes():
    rng = np.random.default_rng(5)
    times = np.linspace(0, 12, 21)  # two passes of up to 16 times
    _, gpu, _ = pair(37, 45, times=times)
    X = torch.as_tensor(rng.standard_normal(gpu.nbasis), device="cuda")
    ad = torch.as_tensor(rng.standard_normal((times.size, 37, 45)), device="cuda")
    F = gpu.operg_batch(times, X)
    G = gpu.operg_transpose_batch(times, ad)
    lhs = float(torch.sum(F * ad))
    rhs = float(torch.dot(X, G))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    assert torch.equal(F, gpu.operg_batch(times, X))
    assert torch.equal(G, gpu.operg_transpose_batch(times, ad))
    for it in (0, 15, 16, 20):
        assert torch.equal(gpu.operg_batch(times[it:it + 1], X)[0], F[it])


def test_shallow_water_configuration_grid():
    """4 scales 50-400 km, 9 directions, 5-day time windows over a 10-day
    window, on the 256 x 256 grid of 8 km cells."""
    lat0 = float(np.degrees(np.arcsin(1e-4 / (2 * 7.2921e-5))))
    times = np.arange(0, 11.0)
    kw = dict(lmin=50.0, lmax=420.0, npsp=3.5, facpsp=3.5, tmin=5.0, tmax=5.0)
    lon, lat = grid(256, 256, lat0=lat0)
    orc = BasisWavelet3dOracle(lon, lat, **kw)
    orc.set_basis(times)
    gpu = Basis_wavelet3d(SimpleNamespace(lon=lon, lat=lat, mask=None), **kw)
    gpu.set_basis(times)
    assert gpu.nbasis == orc.nbasis == 47952
    rng = np.random.default_rng(2)
    X = rng.standard_normal(gpu.nbasis)
    F = gpu.operg_batch(times, X).cpu().numpy()
    for it in (0, 5, 10):
        assert rel(F[it], orc.operg(it, X)) < 1e-12
    ad = rng.standard_normal((2, 256, 256))
    got = gpu.operg_transpose_batch(times[[2, 7]], ad).cpu().numpy()
    ref = orc.operg_transpose(2, ad[0]) + orc.operg_transpose(7, ad[1])
    assert rel(got, ref) < 1e-12
