"""Ragged groups: problems with their own grid sizes, gravity-wave speeds,
windows, observations and checkpoint schedules, advanced by shared launches.

Bar: every member's J and G equal those of its own single-problem engine
(full tape, GEMM configuration 0) bit for bit -- the ragged kernels compute a
member exactly as its own plan does.
"""
import numpy as np
import pytest
import torch

from vardyn_amd import synthetic as syn
from vardyn_amd.fourdvar import FourDVar
from vardyn_amd.problem import build_problem
from vardyn_amd.ragged import RaggedFourDVar

pytestmark = pytest.mark.gpu


def cases(scheme):
    # padded half sizes 32 / 64 / 96 / 128: all three K-split classes of the GEMM
    return [syn.QgCase("r0", 66, 130, 10e3, 2.0, 1200.0, 30, n_tracks=4, track_km=7.0,
                       checkpoint_steps=6, time_scheme=scheme, seed=41),
            syn.QgCase("r1", 194, 98, 10e3, 3.1, 1200.0, 17, n_tracks=6, track_km=7.0,
                       checkpoint_steps=5, time_scheme=scheme, seed=42),
            syn.QgCase("r2", 258, 66, 10e3, 2.5, 1200.0, 41, n_tracks=3, track_km=7.0,
                       checkpoint_steps=7, time_scheme=scheme, seed=43),
            syn.QgCase("r3", 40, 40, 10e3, 1.7, 1200.0, 23, n_tracks=2, track_km=7.0,
                       checkpoint_steps=4, time_scheme=scheme, seed=44)]


def standalone(pb):
    c = pb.case
    return FourDVar(pb.plan, pb.H, pb.checkpoints, pb.T, control="ic", sigma_b=c.sigma_b,
                    tape="full", dtbasis=c.checkpoint_steps)


@pytest.mark.parametrize("scheme", ["Euler", "rk2", "rk3"])
def test_ragged_equals_single_problem_engines(scheme, monkeypatch):
    monkeypatch.setenv("VARDYN_GEMM_CONFIG", "1")
    pbs = [build_problem(c, precision="float64", with_engine=False) for c in cases(scheme)]
    eng = RaggedFourDVar(pbs)
    rng = np.random.default_rng(5)
    X = torch.as_tensor(0.05 * rng.standard_normal(eng.X.numel()), device="cuda")
    J, G = eng.run(X)
    J, Gs = J.clone(), [g.clone() for g in eng.split(G)]
    for i, pb in enumerate(pbs):
        ref = standalone(pb)
        Jr, Gr = ref.run(eng.split(X)[i].contiguous())
        assert torch.equal(J[i:i + 1], Jr), f"member {i}: J differs"
        assert torch.equal(Gs[i], Gr[0]), f"member {i}: G differs"


def test_ragged_graph_replay_matches_eager(monkeypatch):
    monkeypatch.setenv("VARDYN_GEMM_CONFIG", "1")
    pbs = [build_problem(c, precision="float64", with_engine=False) for c in cases("rk2")[:3]]
    eng = RaggedFourDVar(pbs)
    rng = np.random.default_rng(6)
    xs = [torch.as_tensor(0.05 * rng.standard_normal(eng.X.numel()), device="cuda")
          for _ in range(2)]
    eager = []
    for x in xs:
        J, G = eng.run(x)
        eager.append((J.clone(), G.clone()))
    eng.capture()
    for _ in range(2):
        for x, (Je, Ge) in zip(xs, eager):
            J, G = eng.evaluate(x)
            torch.cuda.synchronize()
            assert torch.equal(J, Je) and torch.equal(G, Ge)


def test_group_segment_start_end_equal_plan_calls():
    """vd_group_prep / finish / adj_init / adj_final on a group of uneven
    members (land, an MDT, per-member c) equal the members' own plan calls bit
    for bit."""
    from vardyn_amd import _lib
    from vardyn_amd.qg import QgPlan, build_mask, current_stream_ptr, ptr
    from vardyn_amd.ragged import QgGroup
    rng = np.random.default_rng(7)
    shapes = [(40, 72), (66, 34), (24, 52)]  # even interiors: folded sine transforms
    plans = []
    for k, (ny, nx) in enumerate(shapes):
        ssh = np.zeros((ny, nx))
        ssh[ny // 2, nx // 3] = np.nan
        mdt = 0.05 * rng.standard_normal((ny, nx)) if k == 1 else None
        plans.append(QgPlan(ny, nx, 8e3, 9e3, 900.0, 2.0 + 0.4 * k, 1e-4, time_scheme="rk2",
                            mask=build_mask(ny, nx, ssh), mdt=mdt))
    g = QgGroup(plans)
    tot = g.offsets[-1]
    dev = plans[0].device

    def flat(scale=1.0):
        return scale * torch.randn(tot, dtype=torch.float64, device=dev)

    def view(t, b):
        ny, nx = shapes[b]
        return t[g.offsets[b]:g.offsets[b + 1]].view(ny, nx)

    h0, hb, adh1, lq_in, lh_in = flat(0.1), flat(0.02), flat(), flat(), flat()
    outs = [torch.full((tot,), 7.0, dtype=torch.float64, device=dev) for _ in range(8)]
    h, q, hbm, qb, h1, lh, lq, adh0 = outs
    g.prep(g.n, ptr(h0), ptr(hb), ptr(h), ptr(q), ptr(hbm), ptr(qb))
    g.finish(g.n, ptr(h), ptr(h1))
    g.adj_init(g.n, ptr(adh1), ptr(lh), ptr(lq))
    g.adj_final(g.n, ptr(h), ptr(lh_in), ptr(lq_in), ptr(adh0))
    s = current_stream_ptr(dev)
    for b, p in enumerate(plans):
        e = [torch.empty(shapes[b], dtype=torch.float64, device=dev) for _ in range(6)]
        _lib.call("vd_qg_prep", p._h, ptr(view(h0, b)), ptr(view(hb, b)), ptr(e[0]), ptr(e[1]), s)
        _lib.call("vd_qg_prep", p._h, ptr(view(hb, b)), ptr(view(hb, b)), ptr(e[2]), ptr(e[3]), s)
        _lib.call("vd_qg_finish", p._h, ptr(e[0]), ptr(e[4]), s)
        for got, want in ((h, e[0]), (q, e[1]), (qb, e[3]), (h1, e[4])):
            assert torch.equal(view(got, b).isnan(), want.isnan())
            assert torch.equal(torch.nan_to_num(view(got, b)), torch.nan_to_num(want))
        # hbm = hb (+ mdt): the second prep's h
        assert torch.equal(view(hbm, b), e[2])
        _lib.call("vd_qg_adj_init", p._h, ptr(view(adh1, b)), ptr(e[4]), ptr(e[5]), s)
        assert torch.equal(view(lh, b), e[4]) and torch.equal(view(lq, b), e[5])
        _lib.call("vd_qg_adj_final", p._h, ptr(e[0]), ptr(view(lh_in, b)), ptr(view(lq_in, b)),
                  ptr(e[4]), s)
        assert torch.equal(view(adh0, b), e[4])


def test_banded_group_tables_equal_row_major():
    """A group whose members' stencil fields exceed the 256 MB Infinity Cache
    lays its stencil tiles out in the banded XCD order (with padding entries);
    every member's stage, adjoint stage and segment start equal those of a
    one-member group (row-major table) bit for bit."""
    from vardyn_amd.qg import QgPlan, ptr
    from vardyn_amd.ragged import QgGroup
    ny, nx = 1536, 2048
    plans = [QgPlan(ny, nx, 5e3, 5e3, 900.0, 2.0 + 0.3 * k, 1e-4, time_scheme="rk2",
                    gemm_config=1) for k in range(3)]
    assert 3 * ny * nx * 8 * 4 > 256 << 20 and ny * nx * 8 * 4 < 256 << 20
    g3 = QgGroup(plans)
    tot, m = g3.offsets[-1], ny * nx
    gen = torch.Generator(device="cuda").manual_seed(9)

    def rnd(scale=1.0):
        return scale * torch.randn(tot, dtype=torch.float64, device="cuda", generator=gen)

    h, hb, s0 = rnd(0.1), rnd(0.01), rnd()
    q, hc = torch.empty_like(h), torch.empty_like(h)  # named: the calls must not see freed blocks
    g3.prep(3, ptr(h), ptr(hb), ptr(hc), ptr(q))
    out = [torch.empty_like(h) for _ in range(3)]
    g3.stage(3, ptr(h), ptr(q), ptr(q), None, None, 0, 0, 900.0, ptr(out[0]))
    g3.stage_adj(3, ptr(h), ptr(q), ptr(s0), None, None, 450.0, 0, 0, ptr(s0), None, None,
                 ptr(out[1]), ptr(out[2]))
    for b, p in enumerate(plans):
        g1 = QgGroup([p])
        sl = slice(b * m, (b + 1) * m)
        h1, q1, s1 = h[sl].clone(), torch.empty(m, dtype=torch.float64, device="cuda"), s0[sl].clone()
        hb1, hc1 = hb[sl].clone(), torch.empty_like(h1)
        g1.prep(1, ptr(h1), ptr(hb1), ptr(hc1), ptr(q1))
        assert torch.equal(q1, q[sl])
        o = [torch.empty_like(h1) for _ in range(3)]
        g1.stage(1, ptr(h1), ptr(q1), ptr(q1), None, None, 0, 0, 900.0, ptr(o[0]))
        g1.stage_adj(1, ptr(h1), ptr(q1), ptr(s1), None, None, 450.0, 0, 0, ptr(s1), None, None,
                     ptr(o[1]), ptr(o[2]))
        for k in range(3):
            assert torch.equal(o[k], out[k][sl]), (b, k)
        del g1


@pytest.mark.parametrize("n", [30, 44])
def test_small_member_fused_inversion(n):
    """Groups of equal-size small members (interiors up to 45 points: the tiles
    of QgmWithTiles) invert each member in one workgroup; forward (with q_b and
    h_b) and accumulating inversions against each member's own plan."""
    from vardyn_amd.qg import QgPlan, ptr
    from vardyn_amd.ragged import QgGroup
    ny = nx = n + 2
    plans = [QgPlan(ny, nx, 10e3, 9e3, 1200.0, 2.0 + 0.25 * k, 1e-4 * (1 + 0.1 * k),
                    time_scheme="rk2") for k in range(5)]
    g = QgGroup(plans)
    tot, m = g.offsets[-1], ny * nx
    gen = torch.Generator(device="cuda").manual_seed(13)
    F, Fs, hb, acc = (torch.randn(tot, dtype=torch.float64, device="cuda", generator=gen)
                      for _ in range(4))
    out = torch.empty_like(F)
    g.inversion(len(plans), ptr(F), ptr(out), Fsub=ptr(Fs), hb=ptr(hb))
    acc2 = acc.clone()
    g.inversion(len(plans), ptr(F), ptr(acc2), accumulate=True)
    for b, p in enumerate(plans):
        sl = slice(b * m, (b + 1) * m)
        want = p.pv2h(F[sl].view(ny, nx), hb[sl].view(ny, nx), Fs[sl].view(ny, nx))
        got = out[sl].view(ny, nx)
        assert torch.linalg.vector_norm(got - want) <= 1e-13 * torch.linalg.vector_norm(want)
        pz = p.pv2h(F[sl].view(ny, nx), None, None)
        wacc = acc[sl].view(ny, nx) + pz
        assert torch.linalg.vector_norm(acc2[sl].view(ny, nx) - wacc) <= 1e-13 * torch.linalg.vector_norm(wacc)


def test_group_autotune_and_forced_adjoint_tiles(monkeypatch):
    """The adjoint-stage tile shapes (64 x 16, 32 x 32) and the autotuned
    choice give the same J and G bit for bit; inapplicable choices are
    rejected."""
    from vardyn_amd._lib import VardynError
    monkeypatch.setenv("VARDYN_GEMM_CONFIG", "1")
    pbs = [build_problem(c, precision="float64", with_engine=False) for c in cases("rk2")]
    eng = RaggedFourDVar(pbs)
    rng = np.random.default_rng(9)
    X = torch.as_tensor(0.05 * rng.standard_normal(eng.X.numel()), device="cuda")
    J0, G0 = (t.clone() for t in eng.run(X))
    for tile in ("32x32", "64x16"):
        eng.group.set_tuning({"split_tiles": [None] * 4, "adjoint_tile": tile,
                              "fused_small_inversion": None})
        J1, G1 = eng.run(X)
        assert torch.equal(J0, J1) and torch.equal(G0, G1), tile
    with pytest.raises(VardynError):
        eng.group.set_tuning({"split_tiles": ["64x32"] * 4, "adjoint_tile": "64x16",
                              "fused_small_inversion": None})
    with pytest.raises(VardynError):
        eng.group.set_tuning({"split_tiles": [None] * 4, "adjoint_tile": "64x16",
                              "fused_small_inversion": True})
    tuned = RaggedFourDVar(pbs, autotune=True)
    assert tuned.tuning["split_tiles"] == [None] * 4 and tuned.tuning["fused_small_inversion"] is None
    J2, G2 = tuned.run(X)
    assert torch.equal(J0, J2) and torch.equal(G0, G2)


def test_small_member_group_tuning():
    """Equal-size small members: the fused inversion and the GEMM stages agree
    within rounding, and the autotuned group picks one of them."""
    from vardyn_amd.qg import QgPlan, ptr
    from vardyn_amd.ragged import QgGroup
    ny = nx = 34
    plans = [QgPlan(ny, nx, 10e3, 9e3, 1200.0, 2.0 + 0.25 * k, 1e-4, time_scheme="rk2") for k in range(6)]
    g = QgGroup(plans)
    gen = torch.Generator(device="cuda").manual_seed(21)
    F, hb = (torch.randn(g.offsets[-1], dtype=torch.float64, device="cuda", generator=gen) for _ in range(2))
    outs = []
    for fused in (True, False):
        g.set_tuning({"split_tiles": [None] * 4, "adjoint_tile": "32x32", "fused_small_inversion": fused})
        out = torch.empty_like(F)
        g.inversion(len(plans), ptr(F), ptr(out), hb=ptr(hb))
        outs.append(out)
    torch.cuda.synchronize()
    assert torch.linalg.vector_norm(outs[0] - outs[1]) <= 1e-13 * torch.linalg.vector_norm(outs[1])
    t = g.autotune(reps=2)
    assert t["fused_small_inversion"] in (True, False) and t["split_tiles"] == [None] * 4
