"""The GPU shallow-water model (vardyn_amd/sw.py, vd_sw_* in vd_sw.hip) against
the FP64 oracle (oracle/sw.py), through the C ABI.

Bars: forward steps within 1e-12 relative of the oracle (the kernels contract
multiply-adds into FMAs, the oracle does not); the tangent within 1e-11 of
the oracle's complex-step derivative; the adjoint transposes the oracle's
tangent: <M dx, y> = <dx, M* y> within 1e-11 relative for several random
(dx, y); the reference's own adjoint test (adjoint_test_sw) within 1e-10 of
1; the rest state exactly at rest; repeated calls bitwise equal.  Cases: a
3-layer beta-plane basin, partial slip with bottom drag, a basin with an
island and land corners (which exercises the 6-point blend at the 4-point
vorticity faces), one layer, wind stress (a varying x stress, a uniform y
stress, a mixed-layer depth field), Laplacian viscosity and thickness
diffusion (with land and partial slip), and the 256 x 256 three-layer grid of the
benchmark (a few steps, and the adjoint identity over 12 steps with the
GPU's own tangent).
"""
import numpy as np
import pytest
import torch

from oracle.sw import SwOracle
from vardyn_amd._lib import VardynError
from vardyn_amd.sw import SW

pytestmark = pytest.mark.gpu


def make(nx=20, ny=18, nl=3, slip=0.0, drag=0.0, island=False, seed=0, amp=1.0, wind=False, visc=0.0, diff=0.0,
         aspect=1.0):
    rng = np.random.default_rng(seed)
    wkw = dict(visc_coef=visc, diff_coef=diff) if visc or diff else {}
    if wind:  # a varying x stress, a uniform y stress, a mixed-layer depth field
        wr = np.random.default_rng(99)
        wkw = dict(taux=0.1 * (1 + 0.5 * wr.standard_normal((nx - 1, ny))), tauy=-0.05,
                   h_wind=50.0 + 20.0 * wr.random((nx, ny)), **wkw)
    dx, dy = 8e3, 8e3 * aspect
    y = (np.arange(ny + 1) - ny / 2) * dy
    f = np.broadcast_to(1e-4 + 1.6e-11 * y, (nx + 1, ny + 1)).copy()
    H = np.array([500., 1000., 3000.][:nl]); gp = np.array([0.02, 0.02, 0.01][:nl])
    mask = None
    if island:
        mask = np.ones((nx, ny)); mask[6:10, 5:9] = 0; mask[0, 0] = 0; mask[nx - 1, ny - 3:] = 0
    orc = SwOracle(nx, ny, dx, dy, H, gp, f, 120.0, mask=mask, slip_coef=slip, bottom_drag_coef=drag, **wkw)
    param = dict(nx=nx, ny=ny, nl=nl, dx=dx, dy=dy, H=H.reshape(nl, 1, 1), g_prime=gp, f=f, dt=120.0,
                 mask=mask, slip_coef=slip, bottom_drag_coef=drag, taux=0., tauy=0.)
    param.update(wkw)
    model = SW(param)
    m = orc.m
    x = [amp * 0.1 * rng.standard_normal((nl, nx + 1, ny)) * m["u"],
         amp * 0.1 * rng.standard_normal((nl, nx, ny + 1)) * m["v"],
         amp * 0.5 * rng.standard_normal((nl, nx, ny)) * m["h"]]
    return orc, model, x, rng


def rel(a, b):
    a = a.cpu().numpy() if torch.is_tensor(a) else a
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


CASES = {"basin3": {}, "slip_drag": dict(slip=0.5, drag=1e-6), "island": dict(island=True),
         "one_layer": dict(nl=1), "wind": dict(wind=True, island=True, drag=1e-6),
         "diffusion": dict(visc=2e4, diff=1.5e4, island=True, slip=0.5, aspect=1.3)}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("nstep", [1, 5])
def test_step_matches_oracle(case, nstep):
    orc, model, x, _ = make(**CASES[case])
    got = model.step(*x, nstep=nstep)
    ref = orc.step(*x, nstep=nstep)
    torch.cuda.synchronize()
    for g, r in zip(got, ref):
        assert rel(g, r) < 1e-12


@pytest.mark.parametrize("case", list(CASES))
def test_tangent_matches_complex_step(case):
    orc, model, x, rng = make(**CASES[case])
    m = orc.m
    d = [0.01 * rng.standard_normal(a.shape) * mk for a, mk in zip(x, (m["u"], m["v"], m["h"]))]
    got = model.step_tgl(*x, *d, nstep=4)
    ref = orc.step_tgl(*x, *d, nstep=4)
    torch.cuda.synchronize()
    for g, r in zip(got, ref):
        assert rel(g, r) < 1e-11


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("nstep", [1, 4, 7])
def test_adjoint_transposes_the_oracle_tangent(case, nstep):
    orc, model, x, rng = make(**CASES[case], seed=nstep)
    m = orc.m
    for probe in range(3):
        d = [rng.standard_normal(a.shape) * mk for a, mk in zip(x, (m["u"], m["v"], m["h"]))]
        w = [rng.standard_normal(a.shape) for a in x]
        tl = orc.step_tgl(*x, *d, nstep=nstep)
        (ad,) = model.step_adj(*x, *w, nstep=nstep)
        torch.cuda.synchronize()
        lhs = sum(float(np.sum(a * b)) for a, b in zip(tl, w))
        rhs = sum(float(np.sum(a * b.cpu().numpy())) for a, b in zip(d, ad[:3]))
        assert abs(lhs - rhs) <= 1e-11 * abs(lhs), (probe, lhs, rhs)
        assert ad[3] is None and ad[4] is None


def test_reference_adjoint_test_and_determinism():
    _, model, x, rng = make(island=True)
    assert model.adjoint_test_sw(nstep=6, seed=3) == pytest.approx(1.0, abs=1e-10)
    w = [rng.standard_normal(a.shape) for a in x]
    a1 = [t.clone() for t in model.step_adj(*x, *w, nstep=9)[0][:3]]
    a2 = model.step_adj(*x, *w, nstep=9)[0][:3]
    torch.cuda.synchronize()
    for p, q in zip(a1, a2):
        assert torch.equal(p, q)


def test_sponge_coefficient_changes_nothing():
    """The reference's sponge fields are zero after construction
    (sw.py:374-377), so a sponge coefficient leaves the steps unchanged."""
    _, model, x, _ = make(island=True)
    p = dict(nx=model.nx, ny=model.ny, nl=model.nl, dx=model.dx, dy=model.dy, H=model.H.reshape(-1, 1, 1),
             g_prime=model.g_prime, f=model.f, dt=model.dt, mask=model.mask, slip_coef=model.slip_coef,
             sponge_coef=0.7)
    sp = SW(p)
    a = model.step(*x, nstep=3)
    b = sp.step(*x, nstep=3)
    torch.cuda.synchronize()
    for p_, q_ in zip(a, b):
        assert torch.equal(p_, q_)
    sp.sponge_u = np.zeros((1, 1, model.nx + 1, model.ny))
    with pytest.raises(VardynError):  # a sponge this model does not apply is refused
        sp.sponge_h = np.ones((1, 1, model.nx, model.ny))


def test_rest_state_and_ensemble_axis():
    _, model, x, _ = make(island=True)
    z = [np.zeros((1,) + a.shape) for a in x]
    out = model.step(*z, nstep=3)
    for o in out:
        assert o.shape[0] == 1 and torch.count_nonzero(o) == 0
    (ad,) = model.step_adj(*z, *[np.ones_like(a) for a in z], nstep=0)
    # zero steps: the cotangent comes back masked (set_input_uvh / get_physical_uvh cancel)
    assert torch.allclose(ad[0][0], model.mask_u.expand_as(ad[0][0]), rtol=1e-15, atol=0)


def test_benchmark_grid_against_oracle():
    orc, model, x, rng = make(nx=256, ny=256, nl=3, seed=11)
    got = model.step(*x, nstep=3)
    ref = orc.step(*x, nstep=3)
    torch.cuda.synchronize()
    for g, r in zip(got, ref):
        assert rel(g, r) < 1e-12
    # 12 steps: the adjoint against the GPU tangent
    m = orc.m
    d = [torch.as_tensor(rng.standard_normal(a.shape) * mk, device="cuda") for a, mk in
         zip(x, (m["u"], m["v"], m["h"]))]
    w = [torch.as_tensor(rng.standard_normal(a.shape), device="cuda") for a in x]
    tl = model.step_tgl(*x, *d, nstep=12)
    (ad,) = model.step_adj(*x, *w, nstep=12)
    lhs = sum(float(torch.sum(a * b)) for a, b in zip(tl, w))
    rhs = sum(float(torch.sum(a * b)) for a, b in zip(d, ad[:3]))
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)


def forcing(x, rng, scale=1e-6):
    return [scale * rng.standard_normal(a.shape) for a in x]


@pytest.mark.parametrize("case", ["basin3", "island"])
def test_forced_step_and_tangent_match_oracle(case):
    orc, model, x, rng = make(**CASES[case], seed=21)
    F = forcing(x, rng)
    got = model.step(*x, nstep=5, Fu=F[0], Fv=F[1], Fh=F[2])
    ref = orc.step(*x, nstep=5, F=F)
    torch.cuda.synchronize()
    for g, r in zip(got, ref):
        assert rel(g, r) < 1e-12
    # only the h forcing (the other two None, as the 4D-Var control sets it)
    got = model.step(*x, nstep=3, Fh=F[2])
    ref = orc.step(*x, nstep=3, F=(np.zeros_like(x[0]), np.zeros_like(x[1]), F[2]))
    for g, r in zip(got, ref):
        assert rel(g, r) < 1e-12
    m = orc.m
    d = [0.01 * rng.standard_normal(a.shape) * mk for a, mk in zip(x, (m["u"], m["v"], m["h"]))]
    dF = forcing(x, rng)
    got = model.step_tgl(*x, *d, nstep=4, Fu=F[0], Fv=F[1], Fh=F[2], dFu=dF[0], dFv=dF[1], dFh=dF[2])
    ref = orc.step_tgl(*x, *d, nstep=4, F=F, dF=dF)
    for g, r in zip(got, ref):
        assert rel(g, r) < 1e-11


@pytest.mark.parametrize("nstep", [1, 6])
def test_forcing_adjoint_transposes_the_oracle_tangent(nstep):
    orc, model, x, rng = make(island=True, seed=30 + nstep)
    m = orc.m
    F = forcing(x, rng)
    for probe in range(2):
        d = [rng.standard_normal(a.shape) * mk for a, mk in zip(x, (m["u"], m["v"], m["h"]))]
        dF = forcing(x, rng, 1.0)
        w = [rng.standard_normal(a.shape) for a in x]
        tl = orc.step_tgl(*x, *d, nstep=nstep, F=F, dF=dF)
        out = model.step_adj_forced(*x, *w, nstep=nstep, Fu=F[0], Fv=F[1], Fh=F[2])
        torch.cuda.synchronize()
        lhs = sum(float(np.sum(a * b)) for a, b in zip(tl, w))
        rhs = sum(float(np.sum(a * b.cpu().numpy())) for a, b in zip(d + dF, out))
        assert abs(lhs - rhs) <= 1e-11 * abs(lhs), (probe, lhs, rhs)
    # the unforced adjoint through the forced entry point equals step_adj
    (a1,) = model.step_adj(*x, *w, nstep=nstep)
    a2 = model.step_adj_forced(*x, *w, nstep=nstep, forcing_adjoint=False)
    for p, q in zip(a1[:3], a2[:3]):
        assert torch.equal(p, q)
