"""The device L-BFGS's host logic on the CPU, pinned against scipy (the
reference's minimiser is scipy's L-BFGS-B, mapping/src/inv.py:387-415).

* ``MoreThuente`` against scipy's own Moré-Thuente search (``DCSRCH``) on the
  six test functions of Moré and Thuente (ACM TOMS 20, 1994, section 5): the
  same trial steps and the same accepted step.
* ``DeviceLbfgs``'s driver, with its two vector primitives replaced by NumPy
  (a test double: the product path runs them as HIP kernels), against
  ``scipy.optimize.minimize(method="L-BFGS-B")`` on Rosenbrock, an
  ill-conditioned quadratic and a batch of both: the same accepted iterates,
  iteration counts and stopping reasons.
"""
import math

import numpy as np
import pytest
import scipy.optimize as opt
import torch
from scipy.optimize._dcsrch import DCSRCH

from vardyn_amd.lbfgs import DeviceLbfgs, MoreThuente


# ------------------------------------------------------------ line search
def phi_tests():
    b1, b2 = 2.0, 0.004

    def f1(a):
        return -a / (a * a + b1), -(b1 - a * a) / (a * a + b1) ** 2

    def f2(a):
        return (a + b2) ** 5 - 2 * (a + b2) ** 4, 5 * (a + b2) ** 4 - 8 * (a + b2) ** 3

    def f3(a, beta=0.01, l=39):
        p0 = 1 - a if a <= 1 - beta else (a - 1 if a >= 1 + beta else (a - 1) ** 2 / (2 * beta) + beta / 2)
        d0 = -1 if a <= 1 - beta else (1 if a >= 1 + beta else (a - 1) / beta)
        c = 2 * (1 - beta) / (l * math.pi)
        return p0 + c * math.sin(l * math.pi * a / 2), d0 + (1 - beta) * math.cos(l * math.pi * a / 2)

    def yanai(b1_, b2_):
        g1 = math.sqrt(1 + b1_ ** 2) - b1_
        g2 = math.sqrt(1 + b2_ ** 2) - b2_

        def f(a):
            s1 = math.sqrt((1 - a) ** 2 + b2_ ** 2)
            s2 = math.sqrt(a ** 2 + b1_ ** 2)
            return g1 * s1 + g2 * s2, -g1 * (1 - a) / s1 + g2 * a / s2
        return f
    return [(f1, 1e-3, 1e-1), (f2, 0.1, 0.1), (f3, 0.1, 0.1), (yanai(1e-3, 1e-3), 1e-3, 1e-3),
            (yanai(1e-2, 1e-3), 1e-3, 1e-3), (yanai(1e-3, 1e-2), 1e-3, 1e-3)]


@pytest.mark.parametrize("k", range(6))
@pytest.mark.parametrize("stp0", [1e-3, 1e-1, 10.0, 1e3])
def test_more_thuente_matches_scipy_dcsrch(k, stp0):
    fun, ftol, gtol = phi_tests()[k]
    f0, g0 = fun(0.0)
    ours = MoreThuente(ftol=ftol, gtol=gtol, xtol=0.1, stpmin=0.0, stpmax=1e10)
    ours.start(f0, g0, stp0)
    # scipy's search as a black box: record the steps it evaluates
    seen = []

    def phi(a):
        seen.append(a)
        return fun(a)[0]
    ref = DCSRCH(phi, lambda a: fun(a)[1], ftol, gtol, 0.1, 0.0, 1e10)
    stp_ref, _, _, task_ref = ref(stp0, phi0=f0, derphi0=g0, maxiter=100)
    steps = []
    task = "FG"
    for _ in range(100):
        steps.append(ours.stp)
        f, g = fun(ours.stp)
        task = ours.next(f, g)
        if task != "FG":
            break
    assert task != "FG"
    np.testing.assert_allclose(steps, seen, rtol=1e-12)
    assert task.split(":")[0] == task_ref.decode().split(":")[0]
    if task == "CONVERGENCE":  # scipy returns no step for a search that ends in a warning
        assert ours.stp == pytest.approx(stp_ref, rel=1e-12)


# ------------------------------------------------------------ driver
class HostOps:
    """NumPy stand-ins for the two device primitives of ``DeviceLbfgs``."""

    @staticmethod
    def install(lb, monkeypatch):
        monkeypatch.setattr(lb, "_dots", lambda pairs: [float(np.dot(a.numpy(), b.numpy()))
                                                        for a, b in pairs])

        def axpby(a, x, b, y):
            y.copy_(a * x + b * y)
        monkeypatch.setattr(lb, "_axpby", axpby)


def rosen(x):
    return float(opt.rosen(x)), opt.rosen_der(x)


def quad_factory(n, seed=3):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q @ np.diag(np.logspace(0, 4, n)) @ Q.T
    bb = rng.standard_normal(n)

    def quad(x):
        return float(0.5 * x @ A @ x - bb @ x), A @ x - bb
    return quad


def run_device(funs, x0s, monkeypatch, maxiter):
    B, n = len(funs), x0s[0].size

    def evaluate(X):
        Xh = X.numpy().reshape(B, n)
        out = [f(Xh[b]) for b, f in enumerate(funs)]
        return (torch.tensor([o[0] for o in out], dtype=torch.float64),
                torch.tensor(np.stack([o[1] for o in out]), dtype=torch.float64))
    lb = DeviceLbfgs(evaluate, B, n, torch.device("cpu"))
    HostOps.install(lb, monkeypatch)
    iters = [[] for _ in range(B)]
    X, info = lb.minimize(np.stack(x0s), maxiter=maxiter,
                          callback=lambda b, x: iters[b].append(x.numpy().copy()))
    return X.numpy(), info, iters


def run_scipy(fun, x0, maxiter):
    its = []
    res = opt.minimize(fun, x0, jac=True, method="L-BFGS-B", options={"maxiter": maxiter},
                       callback=lambda xk: its.append(np.array(xk)))
    return res, its


@pytest.mark.parametrize("name", ["rosen2", "rosen10", "quad30"])
def test_driver_matches_scipy_lbfgsb(name, monkeypatch):
    if name.startswith("rosen"):
        n = int(name[5:])
        fun, x0 = rosen, np.full(n, -1.2) + 0.1 * np.arange(n)
    else:
        fun, x0 = quad_factory(30), np.zeros(30)
    maxiter = 200
    res, its = run_scipy(fun, x0, maxiter)
    X, info, iters = run_device([fun], [x0], monkeypatch, maxiter)
    mine = iters[0]
    # the two codes round differently (two-loop recursion against L-BFGS-B's
    # compact matrices) and both problems amplify that over the iterations
    # (quad30, measured: 3e-13 at 25, 1.6e-8 at 100, 1e-3 at 138), so the
    # iterates are compared over the first 25 and the end by its cost
    k = min(len(its), len(mine), 25)
    for i in range(k):
        np.testing.assert_allclose(mine[i], its[i], rtol=1e-7, atol=1e-9)
    assert info["members"][0]["nit"] == res.nit
    assert info["members"][0]["status"] == str(res.message)
    assert fun(X[0])[0] == pytest.approx(res.fun, rel=1e-5, abs=1e-12)


def test_batch_members_run_their_own_searches(monkeypatch):
    """A batch of two different problems in lockstep: each member ends where it
    would alone (its own line searches, memory and stopping test)."""
    funs = [rosen, quad_factory(4, seed=9)]
    x0s = [np.array([-1.2, 1.0, 0.3, -0.5]), np.zeros(4)]
    X, info, _ = run_device(funs, x0s, monkeypatch, 100)
    for b in range(2):
        Xa, infoa, _ = run_device([funs[b]], [x0s[b]], monkeypatch, 100)
        np.testing.assert_array_equal(X[b], Xa[0])
        assert info["members"][b]["nit"] == infoa["members"][0]["nit"]
        res, _ = run_scipy(funs[b], x0s[b], 100)
        assert info["members"][b]["nit"] == res.nit
