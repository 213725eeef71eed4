"""GPU cost_and_grad against the oracle on small seeded problems.

Bars: FP64 plan -- J and G within 1e-10 relative of the FP64 oracle; the two
tape modes (full tape / checkpoint + recompute) agree bit for bit; FP32 plan
-- J within 1e-4 and G within 1e-3 relative over a 1-day window.
"""
import numpy as np
import pytest
import torch

from oracle.fourdvar import BasisInitialState, BasisOffset, CostOracle, Cov, ModelOracle
from oracle.obsop import ObsopOracle
from oracle.qg import QgmOracle
from vardyn_amd import synthetic as syn
from vardyn_amd.fourdvar import FourDVar
from vardyn_amd.problem import build_problem

pytestmark = pytest.mark.gpu


def small_case(scheme="rk2", n=72):
    return syn.QgCase("t64", 64, 72, 10e3, 2.7, 1200.0, n, n_tracks=6, track_km=7.0,
                      checkpoint_steps=12, time_scheme=scheme, seed=21)


def oracle_cost(pb, control):
    c = pb.case
    qg = QgmOracle(dx=c.dx, dy=c.dx, dt=c.dt, c=c.c, f=c.f0, time_scheme=c.time_scheme,
                   shape=(c.ny, c.nx))
    H = ObsopOracle(pb.grid.lon, pb.grid.lat, pb.grid.DX, pb.grid.DY, pb.T, c.dt, pb.dict_obs,
                    Npix=4)
    if control == "ic":
        basis, B = BasisInitialState((c.ny, c.nx), c.sigma_b), Cov(c.sigma_b)
    else:
        basis, B = BasisOffset((c.ny, c.nx), 0.01), Cov(0.01)
    return CostOracle(ModelOracle(qg), H, basis, pb.checkpoints, pb.T, dtbasis=c.checkpoint_steps,
                      B=B, Xb=np.zeros(basis.nbasis))


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


@pytest.mark.parametrize("scheme", ["Euler", "rk2", "rk3"])
def test_cost_grad_ic_f64(scheme):
    pb = build_problem(small_case(scheme), precision="float64", control="ic")
    X = 0.5 * pb.truth0.ravel()
    J, G = pb.engine.cost_and_grad(X)
    Jr, Gr = oracle_cost(pb, "ic").cost_and_grad(X)
    assert J == pytest.approx(Jr, rel=1e-10)
    assert rel(G, Gr) < 1e-10


def test_cost_grad_offset_f64():
    pb = build_problem(small_case(), precision="float64", control="offset")
    X = np.array([0.03])
    J, G = pb.engine.cost_and_grad(X)
    Jr, Gr = oracle_cost(pb, "offset").cost_and_grad(X)
    assert J == pytest.approx(Jr, rel=1e-10)
    assert G[0] == pytest.approx(Gr[0], rel=1e-9)


def test_full_tape_equals_checkpointing_bitwise():
    pb = build_problem(small_case(), precision="float64", control="ic", tape="checkpoint")
    full = FourDVar(pb.plan, pb.H, pb.checkpoints, pb.T, control="ic", sigma_b=0.1, tape="full",
                    dtbasis=pb.case.checkpoint_steps)
    X = torch.as_tensor(0.3 * pb.truth0.ravel(), device=pb.plan.device)
    J1, G1 = pb.engine.run(X)
    J1, G1 = J1.clone(), G1.clone()
    J2, G2 = full.run(X)
    assert torch.equal(J1, J2) and torch.equal(G1, G2)


def test_cost_grad_ic_f32():
    pb = build_problem(small_case(n=72), precision="float32", control="ic")
    X = 0.5 * pb.truth0.ravel()
    J, G = pb.engine.cost_and_grad(X)
    Jr, Gr = oracle_cost(pb, "ic").cost_and_grad(X)
    assert J == pytest.approx(Jr, rel=1e-4)
    assert rel(G, Gr) < 1e-3


def test_gradient_finite_differences_f64():
    """Centred FD of the GPU cost along seeded unit directions with an eps sweep
    1e-2 .. 1e-6 (tools_4Dvar.py:402-413): the best eps agrees with G.d to 1e-6
    (the cost is piecewise smooth: upwinding switches where uT changes sign)."""
    pb = build_problem(small_case(n=36), precision="float64", control="ic")
    X = 0.5 * pb.truth0.ravel()
    J, G = pb.engine.cost_and_grad(X)
    rng = np.random.default_rng(0)
    for _ in range(4):
        d = rng.standard_normal(X.size)
        d /= np.linalg.norm(d)
        gd = np.dot(G, d)
        errs = []
        for eps in (1e-2, 1e-3, 1e-4, 1e-5, 1e-6):
            fd = (pb.engine.cost_and_grad(X + eps * d)[0] -
                  pb.engine.cost_and_grad(X - eps * d)[0]) / (2 * eps)
            errs.append(abs(fd - gd) / abs(gd))
        assert min(errs) < 1e-6, errs


def test_graph_replay_matches_eager():
    """A captured evaluation replayed several times, with eager work and syncs in
    between, reproduces the eager result bit for bit every time."""
    pb = build_problem(small_case(n=48), precision="float64", control="ic", tape="full")
    eng = pb.engine
    rng = np.random.default_rng(7)
    xs = [torch.as_tensor(0.2 * rng.standard_normal(eng.nctrl), device=pb.plan.device)
          for _ in range(3)]
    eager = []
    for x in xs:
        J, G = eng.run(x)
        eager.append((J.clone(), G.clone()))
    eng.capture()
    for rep in range(2):
        for x, (Je, Ge) in zip(xs, eager):
            J, G = eng.replay(x)
            torch.cuda.synchronize()
            assert torch.equal(J, Je) and torch.equal(G, Ge), f"replay {rep} differs"
            # unrelated eager work between replays
            pb.plan.step(pb.plan.field(pb.truth0), None, 3)
            torch.cuda.synchronize()


@pytest.mark.parametrize("prec", [True, False])
def test_cost_grad_bm_basis_f64(prec):
    """SSH-flux control through the balanced-motion wavelet basis (Basis_bm,
    basis.py:472-936) with its per-coefficient Q as sigma_b: GPU engine
    (batched synthesis / transpose over the basis times) against the oracle's
    cost_and_grad with the oracle basis (operg at each basis time)."""
    from types import SimpleNamespace
    from oracle.basis_bm import BasisBmOracle
    from vardyn_amd.basis_bm import Basis_bm
    c = small_case()
    pb = build_problem(c, precision="float64", control="offset")
    cps = np.asarray(pb.checkpoints)
    bseg = [i for i in range(len(cps) - 1) if cps[i] % c.checkpoint_steps == 0]
    tb = pb.T[cps[bseg]] / 3600 / 24
    lon, lat = np.meshgrid(-60 + 0.09 * np.arange(c.nx), 35 + 0.09 * np.arange(c.ny))
    kw = dict(lmin=60.0, lmax=250.0, flux=True)
    gb = Basis_bm(SimpleNamespace(lon=lon, lat=lat, mask=None), **kw)
    _, Q = gb.set_basis(tb, return_q=True)
    orc = BasisBmOracle(lon, lat, None, **kw)
    orc.set_basis(tb)
    eng = FourDVar(pb.plan, pb.H, cps, pb.T, control="basis", basis=gb, sigma_b=Q, prec=prec,
                   dtbasis=c.checkpoint_steps, tape="full")
    rng = np.random.default_rng(4)
    X = rng.standard_normal(gb.nbasis) * (1.0 if prec else Q) * 30.0
    J, G = eng.cost_and_grad(X)
    qg = QgmOracle(dx=c.dx, dy=c.dx, dt=c.dt, c=c.c, f=c.f0, time_scheme=c.time_scheme,
                   shape=(c.ny, c.nx))
    H = ObsopOracle(pb.grid.lon, pb.grid.lat, pb.grid.DX, pb.grid.DY, pb.T, c.dt, pb.dict_obs,
                    Npix=4)
    ref = CostOracle(ModelOracle(qg), H, orc, cps, pb.T, dtbasis=c.checkpoint_steps, B=Cov(Q),
                     Xb=np.zeros(gb.nbasis), prec=prec)
    Jr, Gr = ref.cost_and_grad(X)
    assert J == pytest.approx(Jr, rel=1e-10)
    assert rel(G, Gr) < 1e-9
    # the same evaluation replayed from a captured graph
    eng.capture()
    J2, G2 = eng.cost_and_grad(X)
    assert J2 == J and np.array_equal(G2, G)
