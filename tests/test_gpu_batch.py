"""Batched plans (several problems on one grid, one launch per kernel) against
independent single-problem plans, member by member.

Bar: bit-for-bit equality -- each member runs the same arithmetic in the same
order whether or not it shares the launch.
"""
import numpy as np
import pytest
import torch

from oracle.qg import QgmOracle
from vardyn_amd import synthetic as syn
from vardyn_amd.fourdvar import FourDVar
from vardyn_amd.problem import build_batch
from vardyn_amd.qg import QgPlan

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["float64", "float32"])
@pytest.mark.parametrize("scheme", ["Euler", "rk2", "rk3"])
def test_batched_step_tgl_adj_equal_single(precision, scheme):
    B, ny, nx, n = 3, 40, 48, 5
    cs = [2.0, 2.7, 3.3]
    rng = np.random.default_rng(0)
    h0 = 0.1 * rng.standard_normal((B, ny, nx))
    hb = 0.01 * rng.standard_normal((B, ny, nx))
    dh = 0.1 * rng.standard_normal((B, ny, nx))
    ay = rng.standard_normal((B, ny, nx))
    bp = QgPlan(ny, nx, 10e3, 12e3, 1200.0, cs, 1e-4, time_scheme=scheme, precision=precision,
                nbatch=B)
    H1 = bp.step(bp.field(h0), bp.field(hb), n)
    TL = bp.step_tgl(bp.field(dh), bp.field(h0), bp.field(hb), n)
    AD = bp.step_adj(bp.field(ay), bp.field(h0), bp.field(hb), n)
    for b in range(B):
        sp = QgPlan(ny, nx, 10e3, 12e3, 1200.0, cs[b], 1e-4, time_scheme=scheme,
                    precision=precision)
        assert torch.equal(H1[b], sp.step(sp.field(h0[b]), sp.field(hb[b]), n))
        assert torch.equal(TL[b], sp.step_tgl(sp.field(dh[b]), sp.field(h0[b]), sp.field(hb[b]), n))
        assert torch.equal(AD[b], sp.step_adj(sp.field(ay[b]), sp.field(h0[b]), sp.field(hb[b]), n))


def test_batched_member_matches_oracle_with_its_own_c():
    B, ny, nx, n = 2, 36, 44, 4
    cs = [1.5, 3.0]
    rng = np.random.default_rng(1)
    h0 = 0.1 * rng.standard_normal((B, ny, nx))
    hb = np.zeros((B, ny, nx))
    bp = QgPlan(ny, nx, 10e3, 10e3, 900.0, cs, 1e-4, time_scheme="rk2", precision="float64",
                nbatch=B)
    H1 = bp.step(bp.field(h0), bp.field(hb), n).cpu().numpy()
    for b in range(B):
        ref = QgmOracle(dx=10e3, dy=10e3, dt=900.0, c=cs[b], f=1e-4, time_scheme="rk2",
                        shape=(ny, nx)).step(h0[b], hb[b], n)
        assert np.linalg.norm(H1[b] - ref) / np.linalg.norm(ref) < 1e-10


def test_batched_cost_grad_equals_members():
    case = syn.QgCase("b", 48, 56, 10e3, 2.7, 1200.0, 36, n_tracks=5, track_km=7.0,
                      checkpoint_steps=9, seed=5)
    pb = build_batch(case, 3, precision="float64", tape="full")
    rng = np.random.default_rng(2)
    X = 0.05 * rng.standard_normal((3, case.ny * case.nx))
    J, G = pb.engine.cost_and_grad(X)
    for b in range(3):
        sp = QgPlan(case.ny, case.nx, case.dx, case.dx, case.dt, case.c, case.f0,
                    time_scheme=case.time_scheme, precision="float64")
        eng = FourDVar(sp, pb.Hs[b], pb.checkpoints, pb.T, control="ic", sigma_b=case.sigma_b,
                       tape="checkpoint", dtbasis=case.checkpoint_steps)
        Jb, Gb = eng.cost_and_grad(X[b])
        assert Jb == J[b]
        assert np.array_equal(Gb, G[b])


def test_problem_set_equals_standalone_engines():
    """Problems of different grids and windows evaluated concurrently on several
    streams give exactly the results of their own engines run alone."""
    from vardyn_amd.problem import build_problem
    from vardyn_amd.problemset import ProblemSet
    cases = [syn.QgCase("a", 40, 56, 10e3, 2.0, 1200.0, 30, n_tracks=4, track_km=7.0,
                        checkpoint_steps=6, seed=31),
             syn.QgCase("b", 64, 48, 10e3, 3.1, 1200.0, 22, n_tracks=6, track_km=7.0,
                        checkpoint_steps=5, seed=32),
             syn.QgCase("c", 36, 36, 10e3, 2.5, 1200.0, 41, n_tracks=3, track_km=7.0,
                        checkpoint_steps=7, seed=33)]
    ps = ProblemSet(cases, n_streams=2, autotune=False).capture()
    rng = np.random.default_rng(3)
    X = [torch.as_tensor(0.05 * rng.standard_normal(c.ny * c.nx), device="cuda") for c in cases]
    out, done_ms = ps.evaluate(X)
    assert len(done_ms) == 3 and all(t > 0 for t in done_ms)
    for c, x, (J, G) in zip(cases, X, out):
        eng = build_problem(c, precision="float64", control="ic", tape="full").engine
        Jr, Gr = eng.run(x)
        assert torch.equal(J, Jr) and torch.equal(G, Gr)


def test_banded_stencil_order_equals_single():
    """A batched plan whose stencil fields exceed the 256 MB Infinity Cache runs
    the stage and adjoint-stage stencils in the banded XCD tile order; its
    members stay bit-identical to single plans (row-major order)."""
    B, ny, nx, n = 3, 1536, 2048, 2
    assert B * ny * nx * 8 * 4 > 256 << 20 and ny * nx * 8 * 6 < 256 << 20
    cs = [2.0, 2.4, 2.9]
    rng = np.random.default_rng(1)
    h0 = 0.1 * rng.standard_normal((B, ny, nx))
    hb = 0.01 * rng.standard_normal((B, ny, nx))
    ay = rng.standard_normal((B, ny, nx))
    bp = QgPlan(ny, nx, 5e3, 5e3, 900.0, cs, 1e-4, time_scheme="rk2", nbatch=B, gemm_config=1)
    H1 = bp.step(bp.field(h0), bp.field(hb), n)
    AD = bp.step_adj(bp.field(ay), bp.field(h0), bp.field(hb), n)
    del bp
    for b in range(B):
        sp = QgPlan(ny, nx, 5e3, 5e3, 900.0, cs[b], 1e-4, time_scheme="rk2", gemm_config=1)
        assert torch.equal(H1[b], sp.step(sp.field(h0[b]), sp.field(hb[b]), n))
        assert torch.equal(AD[b], sp.step_adj(sp.field(ay[b]), sp.field(h0[b]), sp.field(hb[b]), n))
