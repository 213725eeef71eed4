"""``Cov``, ``Variational.cost_and_grad`` and ``grad_test`` of
mapping/src/tools_4Dvar.py (datlas-ocean/VarDyn).

``Variational`` keeps the reference's generic segment loop over any
``M`` / ``H`` / ``Basis`` objects with the reference protocols (this
package's ``Model_qg1l``, ``Obsop_interp_l3``, ``Basis_offset``,
``Basis_ic``); the model, observation and reduction work runs on the GPU
through those objects.  For the QG model with initial-state or offset
control, ``vardyn_amd.fourdvar.FourDVar`` runs the same evaluation with
preallocated device buffers and no per-segment host traffic; both give the
same J and G.
"""
from __future__ import annotations

import numpy as np
import torch


class Cov:
    """Diagonal covariance (tools_4Dvar.py:28-44)."""

    def __init__(self, sigma=None):
        self.sigma = 1 if sigma is None else sigma

    def inv(self, X):
        return 1 / self.sigma ** 2 * X

    def sqr(self, X):
        return self.sigma * X

    def invsqr(self, X):
        return 1 / self.sigma * X


class Variational:
    def __init__(self, M=None, H=None, State=None, R=None, B=None, Basis=None, Xb=None,
                 checkpoints=None, dtbasis=None, prec=False):
        self.M, self.H, self.State = M, H, State
        self.R = R if R is not None else Cov(None)
        self.B = B
        self.Xb = Xb
        self.checkpoints = np.asarray(checkpoints)
        self.prec = prec
        self.dtbasis = int(dtbasis) if dtbasis is not None else 1
        self.basis = Basis

    def cost_and_grad(self, X0):
        """J and G of tools_4Dvar.py:252-400 (host float64 in and out)."""
        X0 = np.asarray(X0, dtype=np.float64)
        State = self.State.copy()
        if self.B is not None:
            if self.prec:
                X = self.B.sqr(X0) + self.Xb
                Jb = X0.dot(X0)
            else:
                X = X0 + self.Xb
                Jb = np.dot(X0, self.B.inv(X0))
        else:
            X = X0 - self.Xb
            Jb = 0.0
        ic = getattr(self.basis, "controls_initial_state", False)
        if ic:
            self.basis.operg(0.0, X, State=State.var)
        T = self.M.T
        State_dict, misfit_dict = {}, {}
        Jo = 0.0
        cps = self.checkpoints
        for i in range(len(cps) - 1):
            t = T[cps[i]]
            nstep = int(cps[i + 1] - cps[i])
            if self.H.is_obs_time(t):
                m = self.H.misfit(t, State)
                misfit_dict[t] = m
                Jo += float(torch.dot(m, self.R.inv(m)))
            if (not ic) and cps[i] % self.dtbasis == 0:
                self.basis.operg(t / 3600 / 24, X, State=State.params)
            State_dict[t] = State.copy()
            self.M.step(t=t, State=State, nstep=nstep)
        t = T[-1]
        State_dict[t] = State.copy()
        if self.H.is_obs_time(t):
            m = self.H.misfit(t, State)
            misfit_dict[t] = m
            Jo += float(torch.dot(m, self.R.inv(m)))
        J = 0.5 * (Jo + Jb)

        gb = (X0 if self.prec else self.B.inv(X0)) if self.B is not None else 0
        adState = self.State.copy(free=True)
        adX = X * 0
        t = T[cps[-1]]
        if self.H.is_obs_time(t):
            self.H.adj(t, adState, State_dict[t], misfit_dict[t])
        for i in reversed(range(len(cps) - 1)):
            nstep = int(cps[i + 1] - cps[i])
            t = T[cps[i]]
            self.M.step_adj(t=t, adState=adState, State=State_dict[t], nstep=nstep)
            if (not ic) and cps[i] % self.dtbasis == 0:
                adX = adX + self.basis.operg_transpose(t=t / 3600 / 24, adState=adState.params)
            if self.H.is_obs_time(t):
                self.H.adj(t, adState, State_dict[t], misfit_dict[t])
        if ic:
            adX = adX + self.basis.operg_transpose(0.0, adState.var)
        if self.prec:
            adX = self.B.sqr(adX)
        return J, adX + gb


def grad_test(J, G, X, n=10):
    """|1 - (J(X + l h) - J(X)) / (l G.h)| for l = 10^0 .. 10^-(n-1) (tools_4Dvar.py:402-413)."""
    np.random.seed(0)
    h = np.random.random(X.size)
    h /= np.linalg.norm(h)
    JX = J(X)
    GX = G(X)
    Gh = h.dot(np.where(np.isnan(GX), 0, GX))
    out = []
    for p in range(n):
        lam = 10.0 ** (-p)
        out.append(abs(1.0 - (J(X + lam * h) - JX) / (lam * Gh)))
    return np.array(out)
