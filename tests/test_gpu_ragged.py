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
