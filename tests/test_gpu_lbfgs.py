"""The device L-BFGS (``inv.minimize_device``) with the GPU 4D-Var engine.

* One 128 x 128, 7-day problem: the accepted iterates of 15 iterations
  against scipy's L-BFGS-B driven by the same engine (``inv.minimize``, the
  reference's minimiser): the first within 1e-12, the cost at every iterate
  within 1e-8 (the two codes round their vector arithmetic differently and
  the iteration amplifies it), the same iteration count and stopping reason.
* A batched engine of three problems: each member's iterates are bit for bit
  those of its own single-problem minimisation (own line searches, memory and
  stopping test), and the cost falls at every accepted iterate.
"""
import numpy as np
import pytest
import torch

from vardyn_amd import synthetic as syn
from vardyn_amd.inv import minimize, minimize_device
from vardyn_amd.problem import build_batch, build_problem

pytestmark = pytest.mark.gpu


def test_device_lbfgs_matches_scipy_on_the_engine():
    case = syn.case_basin128()
    pb = build_problem(case, precision="float64", control="ic", tape="full")
    eng = pb.engine.capture()
    xs = []
    x_ref, info_ref = minimize(eng.cost_and_grad, np.zeros(eng.nctrl), maxiter=15,
                               callback=lambda xk: xs.append(np.array(xk)))
    J_ref = [eng.cost_and_grad(xk)[0] for xk in xs]
    xd = []
    X, info = minimize_device(eng, np.zeros(eng.nctrl), maxiter=15,
                              callback=lambda b, x: xd.append(x.cpu().numpy().copy()))
    mem = info["members"][0]
    assert mem["nit"] == info_ref["nit"] == 15
    assert mem["status"] == info_ref["status"].replace("scipy:", "")
    assert np.linalg.norm(xd[0] - xs[0]) <= 1e-12 * np.linalg.norm(xs[0])
    np.testing.assert_allclose(mem["J"][1:], J_ref, rtol=1e-8)
    assert all(b < a for a, b in zip(mem["J"], mem["J"][1:]))


def test_batched_members_equal_single_minimisations(monkeypatch):
    monkeypatch.setenv("VARDYN_GEMM_CONFIG", "1")  # one tile configuration for both plans
    case = syn.case_basin128()
    n_steps = 144
    pb = build_batch(case, 3, precision="float64", n_steps=n_steps, tape="full")
    eng = pb.engine.capture()
    X, info = minimize_device(eng, np.zeros((3, eng.nctrl)), maxiter=8)
    Xb = X.clone()
    for b in range(3):
        assert all(q < p for p, q in zip(info["members"][b]["J"], info["members"][b]["J"][1:]))
    # member 1 alone: the same problem (its own truth, tracks and noise) in a
    # batch of one is built from the batch's pieces
    from vardyn_amd.fourdvar import FourDVar
    from vardyn_amd.qg import QgPlan
    plan1 = QgPlan(case.ny, case.nx, case.dx, case.dx, case.dt, case.c, case.f0, g=case.g,
                   time_scheme=case.time_scheme, precision="float64")
    one = FourDVar(plan1, pb.Hs[1], pb.checkpoints, pb.T, control="ic", sigma_b=case.sigma_b,
                   tape="full", dtbasis=case.checkpoint_steps).capture()
    X1, info1 = minimize_device(one, np.zeros(one.nctrl), maxiter=8)
    assert info1["members"][0]["nit"] == info["members"][1]["nit"]
    assert info1["members"][0]["J"] == info["members"][1]["J"]
    assert torch.equal(X1[0], Xb[1])
