"""Split-product sine GEMMs on int8 MFMA (plans with dst_split = n digits).

Bars: the Helmholtz inversion agrees with the FP64 oracle to the tolerance
written per digit count (the scheme's error is about 2^(5-8n) of the largest
line products: 4e-13 relative for n = 6 on the 510 x 510 transform); every
tile configuration, and a batched plan, give bit-identical results (the
digit sums are exact integers and the FP64 combination runs in a fixed
order); the model step, its tangent and adjoint and the 4D-Var gradient stay
within the written tolerances of the FP64-MFMA plan.
"""
import numpy as np
import pytest
import torch

from test_gpu_qg_parity import make_pair, rel
from vardyn_amd._lib import VardynError
from vardyn_amd.qg import QgPlan

pytestmark = pytest.mark.gpu

TOL = {5: 1e-8, 6: 1e-10, 7: 1e-12}


def split_plan(orc, ny, nx, ns, **kw):
    return QgPlan(ny, nx, orc.dx, orc.dy, orc.dt, orc.c, float(orc.f), time_scheme=orc.time_scheme,
                  precision="float64", dst_split=ns, **kw)


@pytest.mark.parametrize("ns", [5, 6, 7])
@pytest.mark.parametrize("shape", [(40, 48), (66, 34), (128, 128), (41, 47), (40, 45), (130, 258)])
def test_pv2h_against_oracle(shape, ns):
    ny, nx = shape
    orc, _, _, hb, rng = make_pair(ny, nx, "rk2", "float64")
    plan = split_plan(orc, ny, nx, ns)
    q = rng.standard_normal((ny, nx)) * 1e-5
    qb = orc.h2pv(hb, hb)
    h_ref = orc.pv2h(q, hb, qb)
    h = plan.pv2h(plan.field(q), plan.field(hb), plan.field(qb))
    torch.cuda.synchronize()
    assert rel(h.cpu().numpy(), h_ref) < TOL[ns]


@pytest.mark.parametrize("shape", [(130, 258), (258, 258), (257, 259)])
def test_tile_configs_bit_identical(shape):
    ny, nx = shape
    orc, _, _, hb, rng = make_pair(ny, nx, "rk2", "float64")
    q = rng.standard_normal((ny, nx)) * 1e-5
    qb = orc.h2pv(hb, hb)
    outs = []
    for cfg in range(0, 5):
        plan = split_plan(orc, ny, nx, 6, gemm_config=cfg)
        outs.append(plan.pv2h(plan.field(q), plan.field(hb), plan.field(qb)).clone())
    torch.cuda.synchronize()
    for cfg, h in enumerate(outs[1:], start=1):
        assert torch.equal(h, outs[0]), f"configuration {cfg} differs"
    assert rel(outs[0].cpu().numpy(), orc.pv2h(q, hb, qb)) < TOL[6]


def test_batched_plan_bit_identical_to_single_plans():
    ny, nx, nb = 130, 98, 3
    orc, _, _, hb, rng = make_pair(ny, nx, "rk2", "float64")
    cs = np.array([2.0, 2.7, 3.5])
    bplan = QgPlan(ny, nx, orc.dx, orc.dy, orc.dt, cs, float(orc.f), precision="float64",
                   nbatch=nb, dst_split=6)
    q = rng.standard_normal((nb, ny, nx)) * 1e-5
    hbs = 0.01 * rng.standard_normal((nb, ny, nx))
    qbs = np.zeros_like(q)
    hbat = bplan.pv2h(bplan.field(q), bplan.field(hbs), bplan.field(qbs)).clone()
    for b in range(nb):
        plan = QgPlan(ny, nx, orc.dx, orc.dy, orc.dt, cs[b], float(orc.f), precision="float64",
                      dst_split=6)
        h = plan.pv2h(plan.field(q[b]), plan.field(hbs[b]), plan.field(qbs[b]))
        torch.cuda.synchronize()
        assert torch.equal(h, hbat[b]), f"member {b} differs"


@pytest.mark.parametrize("scheme", ["Euler", "rk2", "rk3"])
def test_step_tangent_adjoint_near_f64(scheme):
    ny, nx, nstep = 66, 98, 12
    orc, plan64, h0, hb, rng = make_pair(ny, nx, scheme, "float64", land=True, beta=True, seed=5)
    _, sp, _, _, _ = make_pair(ny, nx, scheme, "float64", land=True, beta=True, seed=5, dst_split=6)
    H0, HB = plan64.field(h0), plan64.field(hb)
    h1 = sp.step(H0, HB, nstep)
    h1_64 = plan64.step(H0, HB, nstep)
    torch.cuda.synchronize()
    assert rel(h1.cpu().numpy(), h1_64.cpu().numpy()) < 1e-9
    dh = plan64.field(np.nan_to_num(rng.standard_normal((ny, nx)) * 1e-3))
    ad = plan64.field(rng.standard_normal((ny, nx)))
    tl = sp.step_tgl(dh, H0, HB, nstep)
    adj = sp.step_adj(ad, H0, HB, nstep)
    tl64 = plan64.step_tgl(dh, H0, HB, nstep)
    adj64 = plan64.step_adj(ad, H0, HB, nstep)
    torch.cuda.synchronize()
    assert rel(tl.cpu().numpy(), tl64.cpu().numpy()) < 1e-9
    assert rel(adj.cpu().numpy(), adj64.cpu().numpy()) < 1e-9
    lhs = float(torch.nansum(tl * ad))
    rhs = float(torch.nansum(dh * adj))
    assert abs(lhs - rhs) <= 1e-9 * max(abs(lhs), abs(rhs))


def test_gradient_near_f64():
    from vardyn_amd import synthetic as syn
    from vardyn_amd.problem import build_problem
    case = syn.QgCase("split", 130, 130, 10e3, 2.5, 1200.0, 96, n_tracks=6, track_km=7.0,
                      checkpoint_steps=24, time_scheme="rk2", seed=11)
    pb = build_problem(case, precision="float64")
    sp = build_problem(case, precision="float64", dst_split=6, truth_plan=pb.plan)
    rng = np.random.default_rng(3)
    X = torch.as_tensor(0.05 * rng.standard_normal(pb.engine.X.numel()), device="cuda")
    J, G = pb.engine.run(X)
    J, G = J.clone(), G.clone()
    Js, Gs = sp.engine.run(X)
    torch.cuda.synchronize()
    assert abs(float(Js[0] - J[0])) <= 1e-10 * abs(float(J[0]))
    assert float(torch.linalg.norm(Gs - G) / torch.linalg.norm(G)) < 1e-9


def test_ragged_group_equals_single_plans():
    """A ragged group of split plans (own grids, windows, schedules) gives every
    member the J and G of its own single-problem engine, bit for bit."""
    from test_gpu_ragged import cases, standalone
    from vardyn_amd.problem import build_problem
    from vardyn_amd.ragged import RaggedFourDVar
    pbs = [build_problem(c, precision="float64", with_engine=False, dst_split=6) for c in cases("rk2")]
    eng = RaggedFourDVar(pbs)
    rng = np.random.default_rng(7)
    X = torch.as_tensor(0.05 * rng.standard_normal(eng.X.numel()), device="cuda")
    J, G = eng.run(X)
    J, Gs = J.clone(), [g.clone() for g in eng.split(G)]
    for i, pb in enumerate(pbs):
        Jr, Gr = standalone(pb).run(eng.split(X)[i].contiguous())
        assert torch.equal(J[i:i + 1], Jr), f"member {i}: J differs"
        assert torch.equal(Gs[i], Gr[0]), f"member {i}: G differs"


@pytest.mark.parametrize("shape", ["64x32", "32x32"])
def test_ragged_group_split_tile_shapes_bit_identical(shape):
    """Both split tile shapes of a ragged group (64 sine lines on zero-padded
    digit planes, or 32) and the autotuned choice give every member the same
    J and G, bit for bit."""
    from test_gpu_ragged import cases
    from vardyn_amd.problem import build_problem
    from vardyn_amd.ragged import RaggedFourDVar
    pbs = [build_problem(c, precision="float64", with_engine=False, dst_split=6) for c in cases("rk2")]
    eng = RaggedFourDVar(pbs)
    rng = np.random.default_rng(8)
    X = torch.as_tensor(0.05 * rng.standard_normal(eng.X.numel()), device="cuda")
    J0, G0 = (t.clone() for t in eng.run(X))
    eng.group.set_tuning({"split_tiles": [shape] * 4, "adjoint_tile": "32x32",
                          "fused_small_inversion": None})
    J1, G1 = eng.run(X)
    assert torch.equal(J0, J1) and torch.equal(G0, G1)
    tuned = RaggedFourDVar(pbs, autotune=True)
    assert all(t in ("64x32", "32x32") for t in tuned.tuning["split_tiles"])
    J2, G2 = tuned.run(X)
    assert torch.equal(J0, J2) and torch.equal(G0, G2)


def test_group_rejects_mixed_split():
    from vardyn_amd.ragged import QgGroup
    orc, plan64, _, _, _ = make_pair(40, 48, "rk2", "float64")
    with pytest.raises(VardynError):
        QgGroup([split_plan(orc, 40, 48, 6), plan64])


# ---------------------------------------------------------------- against the oracle
# The headline batched path runs 5 digits; these compare the split plans with the
# FP64 oracle directly (not with the FP64-MFMA plan).  Bars per digit count: the
# model step, tangent and adjoint over 12 steps within 5e-8 / 5e-10 relative
# (5 / 6 digits: the per-product error of about 2^(5-8n) of the largest line
# products, through 36 inversions), the 4D-Var cost within 1e-9 / 1e-11 and the
# gradient within 5e-8 / 5e-10.
ORACLE_TOL = {5: (5e-8, 1e-9), 6: (5e-10, 1e-11)}


@pytest.mark.parametrize("ns", [5, 6])
@pytest.mark.parametrize("scheme", ["Euler", "rk2", "rk3"])
def test_step_tangent_adjoint_against_oracle(scheme, ns):
    ny, nx, nstep = 66, 98, 12
    orc, sp, h0, hb, rng = make_pair(ny, nx, scheme, "float64", land=True, beta=True, seed=5,
                                     dst_split=ns)
    H0, HB = sp.field(h0), sp.field(hb)
    tol = ORACLE_TOL[ns][0]
    h1 = sp.step(H0, HB, nstep)
    torch.cuda.synchronize()
    assert rel(h1.cpu().numpy(), orc.step(h0, hb, nstep)) < tol
    dhn = np.nan_to_num(rng.standard_normal((ny, nx)) * 1e-3)
    adn = rng.standard_normal((ny, nx))
    tl = sp.step_tgl(sp.field(dhn), H0, HB, nstep)
    adj = sp.step_adj(sp.field(adn), H0, HB, nstep)
    torch.cuda.synchronize()
    tl_ref, _ = orc.step_tgl(dhn, h0, hb, nstep)
    ad_ref = orc.step_adj(adn, h0, hb, nstep)
    ok = ~np.isnan(h0)
    tl_h = tl.cpu().numpy()
    assert np.linalg.norm((tl_h - tl_ref)[ok]) / np.linalg.norm(tl_ref[ok]) < tol
    adj_h = adj.cpu().numpy()
    assert np.linalg.norm((adj_h - ad_ref)[ok]) / np.linalg.norm(ad_ref[ok]) < tol


@pytest.mark.parametrize("ns", [5, 6])
def test_gradient_against_oracle(ns):
    from test_gpu_fourdvar import oracle_cost
    from vardyn_amd import synthetic as syn
    from vardyn_amd.problem import build_problem
    case = syn.QgCase("split", 130, 130, 10e3, 2.5, 1200.0, 96, n_tracks=6, track_km=7.0,
                      checkpoint_steps=24, time_scheme="rk2", seed=11)
    pb = build_problem(case, precision="float64", dst_split=ns)
    X = 0.5 * pb.truth0.ravel()
    J, G = pb.engine.cost_and_grad(X)
    Jr, Gr = oracle_cost(pb, "ic").cost_and_grad(X)
    tg, tj = ORACLE_TOL[ns][0], ORACLE_TOL[ns][1]
    assert abs(J - Jr) <= tj * abs(Jr)
    assert np.linalg.norm(G - Gr) / np.linalg.norm(Gr) < tg
