"""Restatement of the reference 4D-Var cost and gradient -- TEST INFRASTRUCTURE ONLY.

Follows datlas-ocean/VarDyn:

* ``checkpoint_schedule``  -- mapping/src/inv.py:144-163
* ``ModelOracle.step``     -- mapping/src/mod.py:867-913 (``Model_qg1l_jax.step``:
  QG segment, then the flux control ``+ nstep*dt/86400 * Fssh``)
* ``ModelOracle.step_adj`` -- mod.py:1018-1076
* ``Cov``                  -- mapping/src/tools_4Dvar.py:28-44
* ``BasisOffset``          -- mapping/src/basis.py:3939-4000 (``Basis_offset``)
* ``CostOracle.cost_and_grad`` -- tools_4Dvar.py:252-400, including the
  reference's use of the raw misfit (not ``R^{-1}`` misfit) as the
  observation adjoint's input (tools_4Dvar.py:349,380).

Initial-state control (``BasisInitialState``) has no reference class: the
reference computes the initial-state adjoint in its reverse sweep and drops
it (tools_4Dvar.py:356-398).  Here the control vector is copied into the
initial SSH and the gradient is that adjoint.

Pinning: the schedule is checked against hand-worked integer cases; the cost
against closed forms on quadratic toy problems (zero model steps) and the
gradient against centred finite differences along seeded directions.  The
flux increment of ``ModelOracle``: from rest one segment gives exactly
``nstep*dt/86400 * F`` for two segment lengths, the increment is additive
after the segment, the flux adjoint is the pre-propagation state adjoint
times the same scale, and (state, flux) -> state passes the dot-product test
(``tests/test_oracle_pins.py``).
"""
from __future__ import annotations

import numpy as np


def checkpoint_schedule(n_timestamps, nstep_check, is_obs_step):
    """Checkpoint step indices (inv.py:144-163).

    ``is_obs_step(i)`` is the reference's ``Obsop.is_obs(timestamps[i])``.
    """
    checkpoints = [0]
    check = 0
    for i in range(n_timestamps - 1):
        if i > 0 and (is_obs_step(i) or check == nstep_check):
            checkpoints.append(i)
            if check == nstep_check:
                check = 0
        check += 1
    checkpoints.append(n_timestamps - 1)
    return np.asarray(checkpoints)


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


class BasisOffset:
    """Spatially constant flux on SSH (basis.py:3939-4000)."""

    controls_initial_state = False

    def __init__(self, shape, sigma_B):
        self.shape = shape
        self.sigma_B = sigma_B
        self.nbasis = 1

    def set_basis(self):
        Q = self.sigma_B * np.ones(1)
        return np.zeros_like(Q), Q

    def operg(self, X):
        return X * np.ones(self.shape)

    def operg_transpose(self, adparams):
        return np.array([np.sum(adparams)])


class BasisInitialState:
    """Initial SSH as the control vector (no reference counterpart)."""

    controls_initial_state = True

    def __init__(self, shape, sigma_B, land=None):
        self.shape = shape
        self.sigma_B = sigma_B
        self.nbasis = int(np.prod(shape))
        self.land = land

    def set_basis(self):
        Q = self.sigma_B * np.ones(self.nbasis)
        return np.zeros_like(Q), Q

    def operg(self, X):
        """Initial SSH; NaN on land like the reference's initial state (mod.py:697-700)."""
        var = np.asarray(X, dtype=np.float64).reshape(self.shape).copy()
        if self.land is not None:
            var[self.land] = np.nan
        return var

    def operg_transpose(self, ad):
        return np.asarray(ad, dtype=np.float64).ravel().copy()


class ModelOracle:
    """``Model_qg1l_jax`` restricted to SSH (no tracers, no open BC)."""

    def __init__(self, qgm, hb=None):
        self.qgm = qgm
        self.dt = qgm.dt
        self.hb = np.zeros((qgm.ny, qgm.nx)) if hb is None else hb

    def step(self, var, params, nstep):
        x1 = self.qgm.step(var, self.hb, nstep)
        return x1 + nstep * self.dt / (3600 * 24) * params

    def step_adj(self, advar, adparams, var, nstep):
        adparams = adparams + nstep * self.dt / (3600 * 24) * advar
        advar = self.qgm.step_adj(advar, var, self.hb, nstep)
        return advar, adparams


class CostOracle:
    """``Variational.cost_and_grad`` (tools_4Dvar.py:252-400) for SSH control."""

    def __init__(self, model, obsop, basis, checkpoints, T, dtbasis, B=None, R=None, Xb=None,
                 prec=False, var0=None):
        self.M = model
        self.H = obsop
        self.basis = basis
        self.checkpoints = np.asarray(checkpoints)
        self.T = np.asarray(T)
        self.dtbasis = dtbasis
        self.B = B
        self.R = R if R is not None else Cov(None)
        self.Xb = Xb
        self.prec = prec
        shape = (model.qgm.ny, model.qgm.nx)
        self.var0 = np.zeros(shape) if var0 is None else var0

    def cost_and_grad(self, X0):
        X0 = np.asarray(X0, dtype=np.float64)
        if self.B is not None:
            if self.prec:
                X = self.B.sqr(X0) + self.Xb
                Jb = X0.dot(X0)
            else:
                X = X0 + self.Xb
                Jb = np.dot(X0, self.B.inv(X0))
        else:
            X = X0 - self.Xb
            Jb = 0
        var = self.var0.copy()
        params = np.zeros_like(var)
        if self.basis.controls_initial_state:
            var = self.basis.operg(X)
        states, misfits = {}, {}
        Jo = 0.0
        cps = self.checkpoints
        for i in range(len(cps) - 1):
            t = self.T[cps[i]]
            nstep = int(cps[i + 1] - cps[i])
            if self.H.is_obs_time(t):
                m = self.H.misfit(t, var)
                misfits[t] = m
                Jo += m.dot(self.R.inv(m))
            if (not self.basis.controls_initial_state) and cps[i] % self.dtbasis == 0:
                # time-dependent bases take the time in days (tools_4Dvar.py:297)
                params = self.basis.operg(t / 3600 / 24, X) \
                    if getattr(self.basis, "time_dependent", False) else self.basis.operg(X)
            states[t] = (var.copy(), params.copy())
            var = self.M.step(var, params, nstep)
        t = self.T[-1]
        states[t] = (var.copy(), params.copy())
        if self.H.is_obs_time(t):
            m = self.H.misfit(t, var)
            misfits[t] = m
            Jo += m.dot(self.R.inv(m))
        J = 0.5 * (Jo + Jb)

        if self.B is not None:
            gb = X0 if self.prec else self.B.inv(X0)
        else:
            gb = 0
        advar = np.zeros_like(self.var0)
        adparams = np.zeros_like(self.var0)
        adX = X * 0
        t = self.T[cps[-1]]
        if self.H.is_obs_time(t):
            advar = advar + self.H.adj(t, states[t][0], misfits[t])
        for i in reversed(range(len(cps) - 1)):
            nstep = int(cps[i + 1] - cps[i])
            t = self.T[cps[i]]
            advar, adparams = self.M.step_adj(advar, adparams, states[t][0], nstep)
            if (not self.basis.controls_initial_state) and cps[i] % self.dtbasis == 0:
                adX = adX + (self.basis.operg_transpose(t / 3600 / 24, adparams)
                             if getattr(self.basis, "time_dependent", False)
                             else self.basis.operg_transpose(adparams))
                adparams = np.zeros_like(adparams)
            if self.H.is_obs_time(t):
                advar = advar + self.H.adj(t, states[t][0], misfits[t])
        if self.basis.controls_initial_state:
            adX = adX + self.basis.operg_transpose(np.nan_to_num(advar))
        if self.prec:
            adX = self.B.sqr(adX)
        G = adX + gb
        return J, G
