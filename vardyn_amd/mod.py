"""``Model_qg1l_jax`` counterpart (mapping/src/mod.py:470-1076, datlas-ocean/VarDyn):
the ``M`` protocol (``init``, ``set_bc``, ``step``, ``step_tgl``, ``step_adj``;
mod.py:81-143) for the SSH-only QG model on the GPU, plus the reference's
``tangent_test`` and ``adjoint_test`` (mod.py:5327-5446).

``step`` propagates ``State.var['ssh']`` over ``nstep`` steps and then adds
the flux control ``nstep*dt/86400 * State.params['ssh']`` (mod.py:911-913);
``step_adj`` accumulates the flux adjoint before the model adjoint
(mod.py:1058-1068).  Boundary fields (``set_bc``) are taken at the closest
stored time (mod.py:820-849); with none set, hb = 0.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .qg import QgPlan, build_mask, current_stream_ptr, ptr


class Model_qg1l:
    def __init__(self, State, dt, c=2.7, f=1e-4, time_scheme="rk2", init_date=0.0,
                 final_date=None, Kdiffus=None, bathymetry_PV_term=None, mdt=None,
                 SSH=None, precision=None):
        self.name_var = {"SSH": "ssh"}
        self.dt = float(dt)
        self.ny, self.nx = State.ny, State.nx
        precision = precision or ("float64" if State.dtype == torch.float64 else "float32")
        dx = float(np.asarray(State.DX)[self.ny // 2, self.nx // 2]) if State.DX is not None else None
        dy = float(np.asarray(State.DY)[self.ny // 2, self.nx // 2]) if State.DY is not None else None
        beta = None
        if hasattr(f, "__len__"):
            farr = np.asarray(f, dtype=np.float64)
            f0 = float(np.nanmean(farr))
            beta = (farr[2:, :] - farr[:-2, :]) / (2 * dy)
        else:
            f0 = float(f)
        c0 = float(np.nanmean(c)) if hasattr(c, "__len__") else float(c)
        self.f0, self.c = f0, c0
        self.mask = build_mask(self.ny, self.nx, SSH)
        fb = None if bathymetry_PV_term is None else f0 * np.asarray(bathymetry_PV_term, float)
        self.plan = QgPlan(self.ny, self.nx, dx, dy, dt, c0, f0, time_scheme=time_scheme,
                           precision=precision, kdiffus=Kdiffus, mask=self.mask, beta=beta,
                           fbathy=fb, mdt=mdt, device=State.device)
        if final_date is not None:
            self.nt = 1 + int((final_date - init_date) // self.dt)
            self.T = np.arange(self.nt) * self.dt
        self.bc = {}
        self._hb0 = torch.zeros((self.ny, self.nx), dtype=self.plan.dtype, device=self.plan.device)
        self._tape = None
        self._tape_n = -1
        for name in self.name_var.values():
            State.var[name] = torch.zeros((self.ny, self.nx), dtype=State.dtype,
                                          device=State.device)
            if SSH is not None:
                State.var[name][torch.as_tensor(np.isnan(SSH), device=State.device)] = float("nan")
            State.params[name] = torch.zeros((self.ny, self.nx), dtype=State.dtype,
                                             device=State.device)

    # ------------------------------------------------------------------
    def init(self, State, t0=0):
        return

    def set_bc(self, time_bc, var_bc):
        """Boundary SSH fields per time (mod.py:800-818); NaN -> 0."""
        for i, t in enumerate(time_bc):
            a = np.array(var_bc["SSH"][i], dtype=np.float64, copy=True)
            a[np.isnan(a)] = 0.0
            self.bc[float(t)] = self.plan.field(a)

    def _apply_bc(self, t):
        if not self.bc:
            return self._hb0
        times = np.array(list(self.bc.keys()))
        return self.bc[float(times[np.argmin(np.abs(times - t))])]

    def _flux_scale(self, nstep):
        return nstep * self.dt / (3600 * 24)

    def _axpy(self, a, x, y):
        prec = _lib.VD_F64 if y.dtype == torch.float64 else _lib.VD_F32
        _lib.call("vd_axpby", prec, y.numel(), float(a), ptr(x), 1.0, ptr(y),
                  current_stream_ptr(self.plan.device))

    def _tape_for(self, nstep):
        if self._tape is None or self._tape_n < nstep:
            self._tape = self.plan.new_tape(nstep)
            self._tape_n = nstep
        return self._tape

    def step(self, State, nstep=1, t=0):
        hb = self._apply_bc(t)
        x0 = State.var["ssh"].to(self.plan.dtype).contiguous()
        x1 = self.plan.step(x0, hb, int(nstep))
        if "ssh" in State.params:
            self._axpy(self._flux_scale(nstep), State.params["ssh"].to(x1.dtype).contiguous(), x1)
        State.var["ssh"] = x1.to(State.dtype)

    def step_tgl(self, dState, State, nstep=1, t=0):
        hb = self._apply_bc(t)
        dx0 = dState.var["ssh"].to(self.plan.dtype).contiguous()
        x0 = State.var["ssh"].to(self.plan.dtype).contiguous()
        dx1 = self.plan.step_tgl(dx0, x0, hb, int(nstep))
        if "ssh" in dState.params:
            self._axpy(self._flux_scale(nstep), dState.params["ssh"].to(dx1.dtype).contiguous(), dx1)
        dState.var["ssh"] = dx1.to(dState.dtype)

    def step_adj(self, adState, State, nstep=1, t=0):
        hb = self._apply_bc(t)
        ad1 = adState.var["ssh"].to(self.plan.dtype).contiguous()
        x0 = State.var["ssh"].to(self.plan.dtype).contiguous()
        if "ssh" in adState.params:
            p = adState.params["ssh"].to(self.plan.dtype).contiguous()
            self._axpy(self._flux_scale(nstep), ad1, p)
            adState.params["ssh"] = p.to(adState.dtype)
        ad0 = self.plan.step_adj(ad1, x0, hb, int(nstep), tape=self._tape_for(int(nstep)))
        adState.var["ssh"] = ad0.to(adState.dtype)


def _masked_vec(s):
    return np.concatenate((s.getvar(vect=True), s.getparams(vect=True)))


def tangent_test(M, State, t0=0, nstep=1, ampl=1, lambdas=range(10)):
    """Taylor test of ``step_tgl`` (mod.py:5327-5365); returns the ratios."""
    State0 = State.random(ampl=ampl)
    dState = State.random(ampl=ampl)
    rng = np.random.default_rng(1)
    for d in (dState.var, dState.params):
        for k in d:
            d[k] = torch.as_tensor(ampl * rng.random((State.ny, State.nx)),
                                   device=State.device).to(State.dtype)
    S0 = State0.copy()
    M.step(t=t0, State=S0, nstep=nstep)
    X2 = S0.getvar(vect=True)
    out = []
    for p in lambdas:
        lam = 10.0 ** (-p)
        S1 = dState.copy()
        S1.scalar(lam)
        S1.Sum(State0)
        M.step(t=t0, State=S1, nstep=nstep)
        X1 = S1.getvar(vect=True)
        dS = dState.copy()
        dS.scalar(lam)
        M.step_tgl(t=t0, dState=dS, State=State0, nstep=nstep)
        dX = dS.getvar(vect=True)
        m = np.isnan(X1 + X2 + dX)
        out.append(np.linalg.norm(X1[~m] - X2[~m] - dX[~m]) / np.linalg.norm(dX[~m]))
    return np.array(out)


def adjoint_test(M, State, t0=0, nstep=1, ampl=1):
    """<M dx, y> / <dx, M* y> with seeds 10/20/30 (mod.py:5367-5446)."""
    def rand_state(seed):
        np.random.seed(seed)
        s = State.copy(free=True)
        for d in (s.var, s.params):
            for k in d:
                d[k] = torch.as_tensor(ampl * np.random.random((State.ny, State.nx)),
                                       device=State.device).to(State.dtype)
        return s
    S0, dS, adS = rand_state(10), rand_state(20), rand_state(30)
    dX0, adX0 = _masked_vec(dS), _masked_vec(adS)
    M.step_tgl(t=t0, dState=dS, State=S0, nstep=nstep)
    dX1 = _masked_vec(dS)
    M.step_adj(t=t0, adState=adS, State=S0, nstep=nstep)
    adX1 = _masked_vec(adS)
    m = np.isnan(adX0 + dX0 + adX1 + dX1)
    ps1 = np.inner(dX1[~m], adX0[~m])
    ps2 = np.inner(dX0[~m], adX1[~m])
    return ps1 / ps2
