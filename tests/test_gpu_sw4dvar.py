"""The GPU shallow-water 4D-Var engine (vardyn_amd/sw4dvar.py) against the
oracle cost (oracle/sw4dvar.py), through the C ABI.

Bars: J within 1e-12 relative of the oracle's cost; the gradient's
directional derivatives <G, d> within 1e-9 relative of the oracle's
complex-step derivatives (exact to rounding) along seeded directions, with
and without preconditioning; repeated evaluations bitwise equal.  The case:
the 3-layer beta-plane configuration on a 40 x 36 grid over 240 steps, 6
nadir tracks, checkpoints and basis times every 60 steps, a 3-scale wavelet
forcing control, evaluated at half the truth's coefficients (a flowing,
forced state); and the same with Laplacian viscosity and thickness diffusion
(add_diffusion / add_h_diffusion, typical coefficients).
"""
import numpy as np
import pytest
import torch

from oracle.basis_wavelet3d import BasisWavelet3dOracle
from oracle.obsop import ObsopOracle
from oracle.sw import SwOracle
from oracle.sw4dvar import SwCostOracle
from vardyn_amd.sw_problem import basis_kwargs, build_sw_problem, small_case

pytestmark = pytest.mark.gpu


def oracle_for(pb, prec=False):
    c = pb.case
    m = pb.model
    sw = SwOracle(m.nx, m.ny, m.dx, m.dy, m.H, m.g_prime, m.f, m.dt, visc_coef=m.visc_coef, diff_coef=m.diff_coef)
    lon, lat = pb.grid.lon.T, pb.grid.lat.T
    H = ObsopOracle(lon, lat, np.full(lon.shape, c.dx), np.full(lon.shape, c.dx), pb.T, c.dt, pb.dict_obs, Npix=4)
    kw = basis_kwargs(c)
    basis = BasisWavelet3dOracle(lon, lat, **kw)
    basis.set_basis(pb.time_basis)
    return SwCostOracle(sw, H, basis, pb.checkpoints, pb.T, c.checkpoint_steps, pb.x0, sigma_b=pb.Q, prec=prec)


@pytest.mark.parametrize("prec,diffusion", [(False, False), (True, False), (False, True)])
def test_cost_and_gradient_against_oracle(prec, diffusion):
    kw = dict(visc_coef=150.0, diff_coef=100.0) if diffusion else {}
    pb = build_sw_problem(small_case(**kw), prec=prec)
    orc = oracle_for(pb, prec=prec)
    assert orc.basis.nbasis == pb.basis.nbasis
    X0 = 0.5 * pb.X_true / (pb.Q if prec else 1.0)
    J, G = pb.engine.run(torch.as_tensor(X0, device="cuda"))
    J, G = float(J[0]), G.cpu().numpy()
    Jo = orc.cost(X0)
    assert abs(J - Jo) <= 1e-12 * abs(Jo), (J, Jo)
    rng = np.random.default_rng(3)
    for k in range(2 if not prec else 1):
        d = rng.standard_normal(G.size)
        dd = orc.directional_derivative(X0, d)
        assert abs(G @ d - dd) <= 1e-9 * abs(dd), (k, G @ d, dd)
    # determinism
    J2, G2 = pb.engine.run(torch.as_tensor(X0, device="cuda"))
    assert float(J2[0]) == J and np.array_equal(G2.cpu().numpy(), G)


def test_gradient_against_gpu_finite_differences():
    """A second, oracle-free check on a longer window: the GPU gradient against
    centred differences of the GPU cost (the SW model is smooth away from the
    velocity-sign switches of the upwind reconstructions)."""
    pb = build_sw_problem(small_case(days=480 * 120.0 / 86400))
    X0 = torch.as_tensor(0.5 * pb.X_true, device="cuda")
    J, G = pb.engine.run(X0)
    G = G.clone()
    rng = np.random.default_rng(9)
    d = torch.as_tensor(rng.standard_normal(G.numel()), device="cuda")
    eps = 1e-3 * float(torch.linalg.norm(X0)) / float(torch.linalg.norm(d))
    jp = float(pb.engine.run(X0 + eps * d)[0][0])
    jm = float(pb.engine.run(X0 - eps * d)[0][0])
    fd = (jp - jm) / (2 * eps)
    assert abs(float(G @ d) - fd) <= 1e-5 * abs(fd)


def test_tape_modes_bitwise_equal():
    """Every step's state kept (stage states recomputed), sqrt-spaced states
    with each run recomputed, and every step with its stage states (nothing
    recomputed) give the same bits."""
    pb = build_sw_problem(small_case(), tape="full")
    X0 = torch.as_tensor(0.5 * pb.X_true, device="cuda")
    J1, G1 = (t.clone() for t in pb.engine.run(X0))
    assert pb.engine.tape_mode == "full" and int(pb.engine.stride.max()) == 1
    from vardyn_amd.sw4dvar import SwFourDVar
    e2 = SwFourDVar(pb.model, pb.H, pb.checkpoints, pb.T, pb.basis, pb.x0, sigma_b=pb.Q,
                    dtbasis=pb.case.checkpoint_steps, tape="checkpoint")
    assert int(e2.stride.max()) > 1
    J2, G2 = e2.run(X0)
    assert torch.equal(J1, J2) and torch.equal(G1, G2)
    e3 = SwFourDVar(pb.model, pb.H, pb.checkpoints, pb.T, pb.basis, pb.x0, sigma_b=pb.Q,
                    dtbasis=pb.case.checkpoint_steps, tape="stages")
    assert int(e3.stride.max()) == 0
    J3, G3 = e3.run(X0)
    assert torch.equal(J1, J3) and torch.equal(G1, G3)
