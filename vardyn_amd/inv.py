"""4D-Var driver pieces of ``Inv_4Dvar`` (mapping/src/inv.py:105-471, datlas-ocean/VarDyn).

* ``checkpoint_schedule`` -- inv.py:144-163: step 0, every observation step
  (exact date match), every ``nstep_check`` steps (the counter resets only on
  its own hits) and the last step.
* ``minimize`` -- inv.py:266-415: scipy L-BFGS-B on the cost/gradient with the
  reference's ``Wrapper`` semantics: best-state tracking, restart from the best
  state when the cost or gradient is non-finite or ``max|g| >
  gradient_max_norm`` (``CrazyGradient``, at most ``max_retries`` times), and
  optional sustained convergence over ``convergence_nit`` iterations
  (``ConvergenceReached``).
"""
from __future__ import annotations

import time

import numpy as np
import scipy.optimize as opt


class CrazyGradient(Exception):
    """Non-finite or exploding gradient (inv.py:32-33)."""


class ConvergenceReached(Exception):
    """ftol/gtol met over ``convergence_nit`` consecutive iterations (inv.py:29-30)."""


def checkpoint_schedule(n_timestamps, nstep_check, is_obs_step):
    """Checkpoint step indices of inv.py:144-163."""
    cps = [0]
    check = 0
    for i in range(n_timestamps - 1):
        if i > 0 and (is_obs_step(i) or check == nstep_check):
            cps.append(i)
            if check == nstep_check:
                check = 0
        check += 1
    cps.append(n_timestamps - 1)
    return np.asarray(cps, dtype=np.int64)


class _Wrapper:
    """The reference's cost/gradient cache and recovery logic (inv.py:282-385)."""

    def __init__(self, fun, x0, gtol, ftol, convergence_nit, gradient_max_norm, log):
        J0, G0 = fun(x0)
        self.fun = fun
        self.cache = {"cost": J0, "grad": G0}
        self.J_list, self.G_list = [], []
        self.it = 1
        self.t = time.time()
        self.gtol_threshold = gtol
        self.ftol_threshold = ftol
        self.convergence_nit = convergence_nit
        self.gradient_max_norm = gradient_max_norm
        self.convergence_count = 0
        self.last_x = None
        self.best_cost = float(J0)
        self.best_x = np.array(x0).copy()
        self.log = log
        self.nfev = 1

    def __call__(self, x, *args):
        cost, grad = self.fun(x)
        self.nfev += 1
        ftol = (cost - self.cache["cost"]) / max(cost, self.cache["cost"], 1)
        self.cache["cost"], self.cache["grad"] = cost, grad
        mean_grad = float(np.mean(np.abs(grad)))
        self.J_list.append(float(cost))
        self.G_list.append(mean_grad)
        if self.log:
            now = time.time()
            self.log(f"* iteration {self.it} computed in {now - self.t:.2E} s, J={cost:.6E}, "
                     f"G={mean_grad:.3E}")
            self.t = now
        self.it += 1
        gnorm = float(np.max(np.abs(grad)))
        if np.isfinite(cost) and np.isfinite(gnorm) and gnorm < self.gradient_max_norm:
            if cost < self.best_cost:
                self.best_cost = float(cost)
                self.best_x = np.array(x).copy()
        if (not np.isfinite(cost)) or (not np.isfinite(gnorm)) or gnorm > self.gradient_max_norm:
            raise CrazyGradient()
        if self.convergence_nit is not None:
            conv = False
            if self.ftol_threshold is not None and abs(ftol) <= self.ftol_threshold:
                conv = True
            if self.gtol_threshold is not None and gnorm <= self.gtol_threshold:
                conv = True
            self.convergence_count = self.convergence_count + 1 if conv else 0
            self.last_x = np.array(x).copy()
            if self.convergence_count >= self.convergence_nit:
                raise ConvergenceReached()
        return cost

    def jac(self, x, *args):
        return self.cache["grad"]


def minimize(fun, x0, maxiter=10, ftol=None, gtol=None, convergence_nit=None,
             gradient_max_norm=1e6, max_retries=3, method="L-BFGS-B", log=None, callback=None):
    """Minimise ``fun(x) -> (J, G)`` like ``Inv_4Dvar`` (inv.py:266-415).

    Returns ``(x, info)`` with ``info`` holding the cost history, the number of
    iterations and of cost evaluations, and how the loop ended.
    """
    x0 = np.asarray(x0, dtype=np.float64)
    options = {"maxiter": maxiter}
    if ftol is not None:
        options["ftol"] = ftol
    if gtol is not None:
        _, g0 = fun(x0 * 0.0)
        options["gtol"] = gtol * np.max(np.abs(g0))
    ftol_thr, gtol_thr = options.get("ftol"), options.get("gtol")
    if convergence_nit is not None:
        options["ftol"] = 0
        options["gtol"] = 0
    w = _Wrapper(fun, x0, gtol_thr, ftol_thr, convergence_nit, gradient_max_norm, log)
    retry = 0
    status = "maxiter"
    nit = 0
    nfev = 0
    while True:
        try:
            res = opt.minimize(w, x0, method=method, jac=w.jac, options=options, callback=callback)
            x = res.x
            nit += res.nit
            status = "scipy:" + str(res.message)
            break
        except ConvergenceReached:
            x = w.last_x
            status = "converged"
            break
        except CrazyGradient:
            retry += 1
            if retry > max_retries:
                x = w.best_x
                status = "max_retries"
                break
            x0 = w.best_x
            nfev += w.nfev
            w = _Wrapper(fun, x0, gtol_thr, ftol_thr, convergence_nit, gradient_max_norm, log)
    nfev += w.nfev
    return x, {"J": w.J_list, "G": w.G_list, "nit": nit, "nfev": nfev, "status": status,
               "retries": retry, "best_cost": w.best_cost}


def minimize_device(engine, X0, maxiter=10, ftol=None, gtol=None, m=10, maxls=20, callback=None):
    """``minimize`` with the iterates, gradients and correction pairs in HBM
    (``vardyn_amd.lbfgs.DeviceLbfgs``) for a ``FourDVar``-like ``engine``
    (``evaluate(X) -> (J, G)`` on float64 device tensors; a batched engine
    minimises each member on its own).  Options as the reference sets them
    (inv.py:281-293): ``ftol`` as scipy's (default 2.22e-9), ``gtol`` relative
    to ``max|g|`` at X = 0.  Returns ``(X, info)``: X (B, nctrl) on the device,
    ``info["members"]`` the per-member cost history, iteration and evaluation
    counts and stopping reason (scipy's wording)."""
    import torch
    from .lbfgs import DeviceLbfgs
    B, n = engine.B, engine.nctrl
    dev = engine.dev
    X0 = torch.as_tensor(X0, dtype=torch.float64, device=dev).reshape(B, n)
    gtol_abs = 1e-5
    if gtol is not None:
        _, g0 = engine.evaluate(torch.zeros_like(X0))
        gtol_abs = gtol * float(torch.max(torch.abs(g0)))
    lb = DeviceLbfgs(engine.evaluate, B, n, dev, m=m, maxls=maxls)
    return lb.minimize(X0, maxiter=maxiter, ftol=2.220446049250313e-09 if ftol is None else ftol,
                       gtol=gtol_abs, callback=callback)
