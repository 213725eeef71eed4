"""Restatement of the 4D-Var cost of the multilayer shallow-water model with a
reduced-basis forcing control -- TEST INFRASTRUCTURE ONLY.

Follows datlas-ocean/VarDyn:

* ``Variational.cost_and_grad``, cost part -- mapping/src/tools_4Dvar.py:252-312
  (background term, misfits at observation times, the basis field set at
  checkpoints that are multiples of dtbasis, the model over each segment);
* ``Model_qgsw`` with nl > 1 -- mapping/src/mod.py:3003-3016, 3144-3210,
  3292-3358: the observed SSH is the sum of the layer thicknesses; the SSH
  parameter F (per day) becomes the top layer's thickness forcing F / 86400;
* the model: ``oracle.sw.SwOracle.step`` with that forcing; the basis:
  ``oracle.basis_wavelet3d``; the observation tables: ``oracle.obsop``.

Only the cost is restated (the gradient is checked through directional
derivatives of this cost by the complex step: every piece accepts complex
controls and states, so J(X + i eps d) gives <grad J, d> to rounding).
Observations are on the model's (nx, ny) h grid (the State grid transposed),
like the GPU engine.

Pinning (tests/test_oracle_sw4dvar.py): J at rest without forcing is the
closed form 1/2 sum((H 0 - y)/err)^2 + 1/2 X.Q^-2.X; with zero model steps
between observations the cost reduces to the basis and observation
operators alone; the complex-step derivative equals centred finite
differences.
"""
from __future__ import annotations

import numpy as np


class SwCostOracle:
    def __init__(self, sw, obsop, basis, checkpoints, T, dtbasis, x0, sigma_b, Xb=None, sigma_r=None,
                 prec=False):
        self.sw = sw
        self.H = obsop
        self.basis = basis
        self.cps = np.asarray(checkpoints)
        self.T = np.asarray(T, dtype=np.float64)
        self.dtbasis = int(dtbasis)
        self.x0 = tuple(np.asarray(a, dtype=np.float64) for a in x0)
        self.sig = np.broadcast_to(np.asarray(sigma_b, dtype=np.float64), (basis.nbasis,))
        self.Xb = np.zeros(basis.nbasis) if Xb is None else np.asarray(Xb, dtype=np.float64)
        self.sigma_r = 1.0 if sigma_r is None else float(sigma_r)
        self.prec = prec
        self.btimes = np.asarray(basis.time, dtype=np.float64)

    def _misfit(self, t, ssh):
        """Obsop_interp_l3_jax.misfit (obsop.py:610-628), complex-safe."""
        i = self.H._idx(t)
        x = ssh.ravel()
        hx = np.zeros(self.H.y_arr.shape[1], dtype=x.dtype)
        np.add.at(hx, self.H.row_arr[i], self.H.data_arr[i] * x[self.H.col_arr[i]])
        m = hx - self.H.y_arr[i]
        return m / self.H.err_arr[i]

    def _basis_index(self, t_days):
        k = int(np.argmin(np.abs(self.btimes - t_days)))
        assert abs(self.btimes[k] - t_days) < 1e-9, "not a basis time"
        return k

    def cost(self, X0):
        X0 = np.asarray(X0)
        if self.prec:
            X = self.sig * X0 + self.Xb
            Jb = X0.dot(X0)
        else:
            X = X0 + self.Xb
            Jb = np.dot(X0, X0 / self.sig ** 2)
        u, v, h = (a.astype(X.dtype) for a in self.x0)
        sw = self.sw
        zu, zv = np.zeros_like(u), np.zeros_like(v)
        Fh = np.zeros_like(h)
        Jo = 0.0
        for i in range(len(self.cps) - 1):
            t = self.T[self.cps[i]]
            if self.H.is_obs_time(t):
                m = self._misfit(t, h.sum(axis=0))
                Jo = Jo + m.dot(m) / self.sigma_r ** 2
            if self.cps[i] % self.dtbasis == 0:
                phi = self.basis.operg(self._basis_index(t / 3600 / 24), X)
                Fh = np.zeros_like(h)
                Fh[0] = phi / 86400.0
            u, v, h = sw.step(u, v, h, nstep=int(self.cps[i + 1] - self.cps[i]), F=(zu, zv, Fh))
        t = self.T[self.cps[-1]]
        if self.H.is_obs_time(t):
            m = self._misfit(t, h.sum(axis=0))
            Jo = Jo + m.dot(m) / self.sigma_r ** 2
        return 0.5 * (Jo + Jb)

    def directional_derivative(self, X0, d, eps=1e-30):
        """<grad J(X0), d> by the complex step."""
        return float(np.imag(self.cost(np.asarray(X0, dtype=complex) + 1j * eps * np.asarray(d))) / eps)
