"""The balanced-motion wavelet basis on the GPU (vd_bm_operg / vd_bm_operg_adj
behind vardyn_amd.basis_bm.Basis_bm) against the FP64 oracle of Basis_bm
(mapping/src/basis.py:472-936).

Bars: the host tables (frequencies, centres, coefficient layout, Q) equal the
oracle's exactly; operg and operg_transpose within 1e-12 relative of the
oracle (trigonometric functions of the device math library against numpy's);
a batch of times equals the single-time calls (synthesis bit for bit, the
transposed sum within 1e-13); the dot-product identity of the GPU pair; two
runs bit-identical (no atomics).
"""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle.basis_bm import BasisBmOracle
from vardyn_amd.basis_bm import Basis_bm

pytestmark = pytest.mark.gpu


def grid(ny, nx, d, land):
    lon, lat = np.meshgrid(-60 + d * np.arange(nx), 35 + d * np.arange(ny))
    mask = None
    if land:
        mask = np.zeros((ny, nx), dtype=bool)
        mask[:ny // 5, :nx // 4] = True
        mask[ny // 2, nx // 2:nx // 2 + 3] = True
    return lon, lat, mask


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


CASES = [dict(ny=28, nx=32, d=0.05, land=False, flux=False, lmin=30.0, lmax=110.0),
         dict(ny=28, nx=32, d=0.05, land=True, flux=True, lmin=30.0, lmax=110.0),
         dict(ny=60, nx=44, d=0.045, land=True, flux=False, lmin=60.0, lmax=250.0)]


def pair(c, times):
    lon, lat, mask = grid(c["ny"], c["nx"], c["d"], c["land"])
    kw = dict(flux=c["flux"], lmin=c["lmin"], lmax=c["lmax"], delta_time_ref=7.3)
    orc = BasisBmOracle(lon, lat, mask, **kw)
    Xb_o, Q_o = orc.set_basis(times, return_q=True)
    gpu = Basis_bm(SimpleNamespace(lon=lon, lat=lat, mask=mask), **kw)
    Xb_g, Q_g = gpu.set_basis(times, return_q=True)
    return orc, gpu, Q_o, Q_g


@pytest.mark.parametrize("c", CASES)
def test_tables_and_operators_against_oracle(c):
    times = np.arange(0.0, 9.0, 0.5)
    orc, gpu, Q_o, Q_g = pair(c, times)
    assert gpu.nbasis == orc.nbasis > 0 and np.array_equal(Q_g, Q_o)
    assert list(gpu.iff_wavebounds) == list(orc.iff_wavebounds)
    for f in range(orc.nf):
        assert np.array_equal(gpu.centres[f][0], orc.ENSLON[f])
        assert np.array_equal(gpu.centres[f][1], orc.ENSLAT[f])
        assert np.array_equal(gpu.enst[f], orc.enst[f]) and gpu.norm_fact[f] == orc.norm_fact[f]
    rng = np.random.default_rng(1)
    X = rng.standard_normal(gpu.nbasis)
    for t in (0.0, 3.5, 8.5):
        assert rel(gpu.operg(t, X), orc.operg(t, X)) < 1e-12
        phi = rng.standard_normal(gpu.shape_phys)
        adState = {"ssh": phi.copy()}
        got = gpu.operg_transpose(t, adState)
        assert rel(got, orc.operg_transpose(t, phi)) < 1e-12
        assert not np.any(adState["ssh"])             # consumed, as in the reference


@pytest.mark.parametrize("c", CASES[1:])
def test_batch_adjoint_and_determinism(c):
    times = np.arange(0.0, 9.0, 0.5)
    _, gpu, _, _ = pair(c, times)
    rng = np.random.default_rng(2)
    X = torch.as_tensor(rng.standard_normal(gpu.nbasis), device="cuda")
    F = gpu.operg_batch(times, X)
    for k in (0, 5, len(times) - 1):
        assert torch.equal(F[k], gpu.operg_batch([times[k]], X)[0])
    assert torch.equal(F, gpu.operg_batch(times, X))
    A = torch.as_tensor(rng.standard_normal((len(times),) + gpu.shape_phys), device="cuda")
    adX = gpu.operg_transpose_batch(times, A)
    assert torch.equal(adX, gpu.operg_transpose_batch(times, A))
    one = sum(gpu.operg_transpose_batch([t], A[k]) for k, t in enumerate(times))
    assert rel(adX.cpu(), one.cpu()) < 1e-13
    lhs = float((F * A).sum())
    rhs = float((X * adX).sum())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_state_protocol():
    c = CASES[0]
    times = np.arange(0.0, 4.0)
    _, gpu, _, _ = pair(c, times)
    X = np.random.default_rng(3).standard_normal(gpu.nbasis)
    State = {"ssh": torch.zeros(gpu.shape_phys, dtype=torch.float64, device="cuda")}
    gpu.operg(1.0, X, State=State)
    assert torch.equal(State["ssh"].cpu(), torch.as_tensor(gpu.operg(1.0, X)))


# ------------------------------------------------- geostrophic velocities (basis.py:822-860)
def uv_pair(ny=31, nx=26, seed=4):
    from oracle.basis_bm import geostrophic_tables
    lon, lat, mask = grid(ny, nx, 0.05, False)
    rng = np.random.default_rng(seed)
    f = 1e-4 + 2e-6 * rng.standard_normal((ny, nx))
    DX = 5e3 * (1 + 0.1 * rng.random((ny, nx)))
    DY = 6e3 * (1 + 0.1 * rng.random((ny, nx)))
    gpu = Basis_bm(SimpleNamespace(lon=lon, lat=lat, mask=mask, f=f, DX=DX, DY=DY), lmin=30.0, lmax=110.0,
                   compute_velocities=True)
    gpu.set_basis(np.arange(0.0, 6.0, 0.5))
    return gpu, geostrophic_tables(f, DX, DY), rng


def test_ssh2uv_and_adjoint_against_oracle():
    """Device velocities and their adjoint equal the oracle's (the reference's
    operation order, no contraction) for three fields at once."""
    from oracle.basis_bm import ssh2uv, ssh2uv_adj
    gpu, tab, rng = uv_pair()
    ny, nx, T = gpu.ny, gpu.nx, 3
    ssh = rng.standard_normal((T, ny, nx))
    u, v = gpu.ssh2uv_batch(torch.as_tensor(ssh))
    adu, adv = rng.standard_normal((T, ny, nx + 1)), rng.standard_normal((T, ny + 1, nx))
    a0 = rng.standard_normal((T, ny, nx))
    ad = gpu.ssh2uv_adj_batch(torch.as_tensor(adu), torch.as_tensor(adv), torch.as_tensor(a0, device="cuda").clone())
    u, v, ad = u.cpu().numpy(), v.cpu().numpy(), ad.cpu().numpy()
    for t in range(T):
        uo, vo = ssh2uv(ssh[t], *tab)
        assert rel(u[t], uo) <= 1e-15 and rel(v[t], vo) <= 1e-15
        assert rel(ad[t], a0[t] + ssh2uv_adj(adu[t], adv[t], *tab)) <= 1e-15


def test_operg_with_velocities():
    """operg with compute_velocities fills the State's u and v with the
    velocities of its SSH; operg_transpose folds the velocity adjoints into
    the SSH adjoint before the transpose and zeroes all three."""
    gpu, tab, rng = uv_pair()
    X = rng.standard_normal(gpu.nbasis)
    ssh, u, v = gpu.operg(1.5, X)
    u2, v2 = gpu.ssh2uv_batch(torch.as_tensor(ssh)[None])
    assert np.array_equal(u, u2[0].cpu().numpy()) and np.array_equal(v, v2[0].cpu().numpy())
    st = {"ssh": np.zeros_like(ssh), "u": None, "v": None}
    gpu.operg(1.5, torch.as_tensor(X, device="cuda"), State=st)
    assert np.array_equal(st["u"].cpu().numpy(), u)
    adssh = rng.standard_normal(ssh.shape)
    adu, adv = rng.standard_normal(u.shape), rng.standard_normal(v.shape)
    ad = {"ssh": adssh.copy(), "u": adu.copy(), "v": adv.copy()}
    adX = gpu.operg_transpose(1.5, ad)
    a = torch.as_tensor(adssh, device="cuda")[None].clone()
    gpu.ssh2uv_adj_batch(torch.as_tensor(adu), torch.as_tensor(adv), a)
    want = gpu.operg_transpose_batch([1.5], a).cpu().numpy()
    assert np.array_equal(adX, want)
    assert not ad["ssh"].any() and not ad["u"].any() and not ad["v"].any()
    # the pair passes the dot-product test
    lhs = np.sum(ssh * adssh) + np.sum(u * adu) + np.sum(v * adv)
    assert abs(lhs - np.dot(X, adX)) <= 1e-12 * abs(lhs)


def test_basis_factory_builds_bm():
    """The reference's factory (basis.py:30-86) with BASIS_BM and
    BASIS_BM_JAX builds the GPU basis with the config's fields."""
    import datetime

    from vardyn_amd.basis import Basis
    lon, lat, mask = grid(20, 24, 0.05, False)
    for sup in ("BASIS_BM", "BASIS_BM_JAX"):
        cfg = SimpleNamespace(BASIS=SimpleNamespace(super=sup, lmin=30.0, lmax=110.0, flux=False,
                                                    name_mod_var="ssh"),
                              EXP=SimpleNamespace(init_date=datetime.datetime(2012, 10, 1)))
        b = Basis(cfg, SimpleNamespace(lon=lon, lat=lat, mask=mask), verbose=False)
        assert isinstance(b, Basis_bm) and b.lmin == 30.0
        assert b.delta_time_ref == (datetime.datetime(2012, 10, 1) - datetime.datetime(1950, 1, 1)).days
