"""Assemble a complete synthetic QG 4D-Var problem on the GPU.

Grid, truth run, observations (sampled from the truth on the model itself),
observation operator tables, checkpoint schedule and the cost/gradient
engine -- what ``Inv_4Dvar`` sets up before minimising (mapping/src/inv.py:105-263).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import synthetic as syn
from .fourdvar import FourDVar
from .inv import checkpoint_schedule
from .obsop import Obsop_interp_l3
from .qg import QgPlan


@dataclass
class Problem:
    case: syn.QgCase
    grid: syn.Grid
    plan: QgPlan
    H: Obsop_interp_l3
    T: np.ndarray
    checkpoints: np.ndarray
    obs: syn.ObsSet
    dict_obs: dict
    truth0: np.ndarray
    engine: FourDVar | None = None


def truth_run(plan: QgPlan, h0, steps, dt):
    """SSH of a free run from ``h0`` at the requested steps (host arrays)."""
    out = {}
    h = plan.field(h0)
    hb = torch.zeros_like(h)
    nxt = torch.empty_like(h)
    s0 = 0
    for s in sorted(steps):
        if s > s0:
            plan.step(h, hb, s - s0, h1=nxt)
            h, nxt = nxt, h
        out[s] = h.double().cpu().numpy()
        s0 = s
    return out


def build_batch(case: syn.QgCase, nbatch, precision="float64", n_steps=None, tape="full",
                device=None, c=None, dst_split=0):
    """``nbatch`` independent problems of ``case`` in one batched plan.

    Member b has its own seeded initial field, track/swath geometry and
    noise; all members share the observation steps, so they share the
    checkpoint schedule.  ``c`` may give one gravity-wave speed per member.
    """
    n_steps = case.n_steps if n_steps is None else int(n_steps)
    grid = syn.make_grid(case.ny, case.nx, case.dx)
    plan = QgPlan(case.ny, case.nx, case.dx, case.dx, case.dt, case.c if c is None else c,
                  case.f0, g=case.g, time_scheme=case.time_scheme, precision=precision,
                  device=device, nbatch=nbatch, dst_split=dst_split)
    truth0 = np.stack([syn.gaussian_random_field(case.ny, case.nx, case.dx, case.wavelength,
                                                 case.rms, case.seed + 7919 * b)
                       for b in range(nbatch)])
    obs = [syn.make_observations(case, grid, n_steps, member=b) for b in range(nbatch)]
    steps = sorted(obs[0].steps)
    fields = truth_run(plan, truth0 if nbatch > 1 else truth0[0], steps, case.dt)
    T = np.arange(n_steps + 1) * case.dt
    Hs, dicts = [], []
    for b in range(nbatch):
        sl = {s: (fields[s][b] if nbatch > 1 else fields[s]) for s in steps}
        syn.sample_truth(obs[b], lambda st, sl=sl: sl, case.sigma_obs, case.seed + 3 + 1009 * b)
        d = {float(T[s]): (e["lon"], e["lat"], e["y"], e["err"]) for s, e in obs[b].steps.items()}
        dicts.append(d)
        Hs.append(Obsop_interp_l3(grid.lon, grid.lat, grid.DX, grid.DY, T, case.dt, d, Npix=4,
                                  device=plan.device).process_obs())
    cps = checkpoint_schedule(n_steps + 1, case.checkpoint_steps, lambda i: Hs[0].is_obs(T[i]))
    pb = Problem(case, grid, plan, Hs[0], T, cps, obs[0], dicts[0], truth0[0])
    pb.Hs, pb.dicts, pb.truth = Hs, dicts, truth0
    pb.engine = FourDVar(plan, Hs if nbatch > 1 else Hs[0], cps, T, control="ic",
                         sigma_b=case.sigma_b, tape=tape, dtbasis=case.checkpoint_steps)
    return pb


def build_problem(case: syn.QgCase, precision="float64", n_steps=None, control="ic",
                  tape="checkpoint", device=None, truth_plan: QgPlan | None = None,
                  with_engine=True, dst_split=0) -> Problem:
    n_steps = case.n_steps if n_steps is None else int(n_steps)
    grid = syn.make_grid(case.ny, case.nx, case.dx)
    plan = QgPlan(case.ny, case.nx, case.dx, case.dx, case.dt, case.c, case.f0, g=case.g,
                  time_scheme=case.time_scheme, precision=precision, device=device,
                  dst_split=dst_split)
    obs = syn.make_observations(case, grid, n_steps)
    truth0 = syn.gaussian_random_field(case.ny, case.nx, case.dx, case.wavelength, case.rms,
                                       case.seed)
    tp = truth_plan if truth_plan is not None else plan
    syn.sample_truth(obs, lambda steps: truth_run(tp, truth0, steps, case.dt), case.sigma_obs,
                     case.seed + 3)
    T = np.arange(n_steps + 1) * case.dt
    dict_obs = {float(T[s]): (e["lon"], e["lat"], e["y"], e["err"]) for s, e in obs.steps.items()}
    H = Obsop_interp_l3(grid.lon, grid.lat, grid.DX, grid.DY, T, case.dt, dict_obs, Npix=4,
                        device=plan.device).process_obs()
    cps = checkpoint_schedule(n_steps + 1, case.checkpoint_steps, lambda i: H.is_obs(T[i]))
    pb = Problem(case, grid, plan, H, T, cps, obs, dict_obs, truth0)
    if with_engine:
        sb = case.sigma_b if control == "ic" else 0.01
        pb.engine = FourDVar(plan, H, cps, T, control=control, sigma_b=sb, tape=tape,
                             dtbasis=case.checkpoint_steps)
    return pb
