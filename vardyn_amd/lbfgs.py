"""L-BFGS minimisation with the control, gradient and correction pairs in HBM.

The reference minimises the 4D-Var cost with scipy's L-BFGS-B (no bounds,
m = 10 correction pairs; mapping/src/inv.py:387-415, through ``Inv_4Dvar``'s
``Wrapper``, inv.py:282-385).  This module runs the same method with every
vector on the device: the inner products in FP64 (``vd_dot_f64``,
deterministic wavefront reductions) and the vector updates with
``vd_axpby``; only scalars (J, directional derivatives, dot products) come
back to the host, where the line search and the stopping tests run.

What it follows, so that its iterates are L-BFGS-B's for unbounded problems:

* Direction: the two-loop recursion with H0 = (s'y / y'y) I of the newest
  pair.  L-BFGS-B's compact matrix B (B0 = theta I, theta = y'y / s'y) is the
  inverse of that H, and without bounds its Cauchy point and subspace
  minimisation give z - x = -B^-1 g, so the steps are the same up to
  rounding.  The first direction is -g.
* Line search: the Moré-Thuente search (Moré and Thuente, ACM TOMS 20,
  1994; MINPACK-2 ``dcsrch`` / ``dcstep``) with L-BFGS-B's settings:
  ftol 1e-3, gtol 0.9, xtol 0.1, stpmin 0, stpmax 1e10; the first trial step
  is 1 / ||d|| on the first iteration and 1 afterwards; at most ``maxls``
  evaluations per search, after which the memory is dropped and the
  iteration restarts from the last accepted point (twice in a row ends it).
* Pair update: skipped when s'y <= eps * (-g_old'd) (L-BFGS-B's test).
* Stopping (scipy's options): ``max|g| <= gtol``; ``(f_old - f) <=
  ftol * max(|f_old|, |f|, 1)``; ``maxiter`` accepted iterations.

A batch of B independent problems (the batched engines) runs B searches in
lockstep: each round evaluates all members at their current trial points
(members that have stopped are evaluated at their final point and ignore
the result), so a round costs one batched evaluation.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib
from ._lib import VardynError
from .qg import current_stream_ptr, ptr

EPS = float(np.finfo(np.float64).eps)
XTRAPL, XTRAPU = 1.1, 4.0


class MoreThuente:
    """Scalar state machine of the Moré-Thuente line search (dcsrch).

    ``start(f0, g0, stp)`` begins a search at the accepted point (value f0,
    directional derivative g0 < 0) with a first trial step; each
    ``next(f, g)`` takes the value and directional derivative at the current
    trial ``self.stp`` and returns "FG" (evaluate at the new ``self.stp``),
    "CONVERGENCE" or "WARNING: ..." (accept ``self.stp``)."""

    def __init__(self, ftol=1e-3, gtol=0.9, xtol=0.1, stpmin=0.0, stpmax=1e10):
        self.ftol, self.gtol, self.xtol = ftol, gtol, xtol
        self.stpmin, self.stpmax = stpmin, stpmax

    def start(self, f0, g0, stp):
        if not g0 < 0:
            raise VardynError("line search needs a descent direction")
        if stp < self.stpmin or stp > self.stpmax:
            raise VardynError("initial step outside [stpmin, stpmax]")
        self.stp = float(stp)
        self.brackt = False
        self.stage = 1
        self.finit, self.ginit = float(f0), float(g0)
        self.gtest = self.ftol * self.ginit
        self.width = self.stpmax - self.stpmin
        self.width1 = self.width / 0.5
        self.stx, self.fx, self.gx = 0.0, self.finit, self.ginit
        self.sty, self.fy, self.gy = 0.0, self.finit, self.ginit
        self.stmin = 0.0
        self.stmax = self.stp + XTRAPU * self.stp
        return "FG"

    def next(self, f, g):
        stp = self.stp
        ftest = self.finit + stp * self.gtest
        if self.stage == 1 and f <= ftest and g >= 0:
            self.stage = 2
        task = None
        if self.brackt and (stp <= self.stmin or stp >= self.stmax):
            task = "WARNING: ROUNDING ERRORS PREVENT PROGRESS"
        if self.brackt and self.stmax - self.stmin <= self.xtol * self.stmax:
            task = "WARNING: XTOL TEST SATISFIED"
        if stp == self.stpmax and f <= ftest and g <= self.gtest:
            task = "WARNING: STP = STPMAX"
        if stp == self.stpmin and (f > ftest or g >= self.gtest):
            task = "WARNING: STP = STPMIN"
        if f <= ftest and abs(g) <= self.gtol * (-self.ginit):
            task = "CONVERGENCE"
        if task is not None:
            return task
        if self.stage == 1 and f <= self.fx and f > ftest:
            # the modified function psi(stp) = f(stp) - f(0) - stp * gtest
            gt = self.gtest
            fm, gm = f - stp * gt, g - gt
            fxm, gxm = self.fx - self.stx * gt, self.gx - gt
            fym, gym = self.fy - self.sty * gt, self.gy - gt
            (self.stx, fxm, gxm, self.sty, fym, gym, self.stp, self.brackt) = dcstep(
                self.stx, fxm, gxm, self.sty, fym, gym, stp, fm, gm, self.brackt,
                self.stmin, self.stmax)
            self.fx, self.gx = fxm + self.stx * gt, gxm + gt
            self.fy, self.gy = fym + self.sty * gt, gym + gt
        else:
            (self.stx, self.fx, self.gx, self.sty, self.fy, self.gy, self.stp, self.brackt) = dcstep(
                self.stx, self.fx, self.gx, self.sty, self.fy, self.gy, stp, f, g, self.brackt,
                self.stmin, self.stmax)
        if self.brackt:
            if abs(self.sty - self.stx) >= 0.66 * self.width1:
                self.stp = self.stx + 0.5 * (self.sty - self.stx)
            self.width1 = self.width
            self.width = abs(self.sty - self.stx)
            self.stmin = min(self.stx, self.sty)
            self.stmax = max(self.stx, self.sty)
        else:
            self.stmin = self.stp + XTRAPL * (self.stp - self.stx)
            self.stmax = self.stp + XTRAPU * (self.stp - self.stx)
        self.stp = min(max(self.stp, self.stpmin), self.stpmax)
        if (self.brackt and (self.stp <= self.stmin or self.stp >= self.stmax)) or \
                (self.brackt and self.stmax - self.stmin <= self.xtol * self.stmax):
            self.stp = self.stx
        return "FG"


def dcstep(stx, fx, dx, sty, fy, dy, stp, fp, dp, brackt, stpmin, stpmax):
    """One safeguarded step of the Moré-Thuente search: a cubic or quadratic
    interpolation step from the interval end points (stx, sty) and the trial
    stp, and the updated interval (the four cases of the paper)."""
    sgnd = dp * math.copysign(1.0, dx)
    if fp > fx:
        # 1: higher value -- the minimum is bracketed
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt((theta / s) ** 2 - (dx / s) * (dp / s))
        if stp < stx:
            gamma = -gamma
        p = (gamma - dx) + theta
        q = ((gamma - dx) + gamma) + dp
        stpc = stx + (p / q) * (stp - stx)
        stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx)
        stpf = stpc if abs(stpc - stx) < abs(stpq - stx) else stpc + (stpq - stpc) / 2.0
        brackt = True
    elif sgnd < 0:
        # 2: lower value, derivatives of opposite sign -- bracketed
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt((theta / s) ** 2 - (dx / s) * (dp / s))
        if stp > stx:
            gamma = -gamma
        p = (gamma - dp) + theta
        q = ((gamma - dp) + gamma) + dx
        stpc = stp + (p / q) * (stx - stp)
        stpq = stp + (dp / (dp - dx)) * (stx - stp)
        stpf = stpc if abs(stpc - stp) > abs(stpq - stp) else stpq
        brackt = True
    elif abs(dp) < abs(dx):
        # 3: lower value, same sign, derivative magnitude decreases
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt(max(0.0, (theta / s) ** 2 - (dx / s) * (dp / s)))
        if stp > stx:
            gamma = -gamma
        p = (gamma - dp) + theta
        q = (gamma + (dx - dp)) + gamma
        r = p / q
        if r < 0 and gamma != 0:
            stpc = stp + r * (stx - stp)
        elif stp > stx:
            stpc = stpmax
        else:
            stpc = stpmin
        stpq = stp + (dp / (dp - dx)) * (stx - stp)
        if brackt:
            stpf = stpc if abs(stpc - stp) < abs(stpq - stp) else stpq
            if stp > stx:
                stpf = min(stp + 0.66 * (sty - stp), stpf)
            else:
                stpf = max(stp + 0.66 * (sty - stp), stpf)
        else:
            stpf = stpc if abs(stpc - stp) > abs(stpq - stp) else stpq
            stpf = max(stpmin, min(stpmax, stpf))
    else:
        # 4: lower value, same sign, derivative magnitude does not decrease
        if brackt:
            theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp
            s = max(abs(theta), abs(dy), abs(dp))
            gamma = s * math.sqrt((theta / s) ** 2 - (dy / s) * (dp / s))
            if stp > sty:
                gamma = -gamma
            p = (gamma - dp) + theta
            q = ((gamma - dp) + gamma) + dy
            stpf = stp + (p / q) * (sty - stp)
        elif stp > stx:
            stpf = stpmax
        else:
            stpf = stpmin
    if fp > fx:
        sty, fy, dy = stp, fp, dp
    else:
        if sgnd < 0:
            sty, fy, dy = stx, fx, dx
        stx, fx, dx = stp, fp, dp
    return stx, fx, dx, sty, fy, dy, stpf, brackt


class _Member:
    """Per-problem state of the batched minimiser (host scalars)."""

    def __init__(self, m):
        self.m = m
        self.rho = []          # 1 / s'y of the stored pairs, oldest first
        self.slots = []        # memory slot of each stored pair, oldest first
        self.ls = MoreThuente()
        self.done = False
        self.status = "running"
        self.nit = 0
        self.nfev = 0
        self.J_hist = []
        self.fails = 0


class DeviceLbfgs:
    """L-BFGS over ``B`` independent problems of ``n`` controls each, with all
    vectors in device memory.

    ``evaluate(X)`` takes a float64 device tensor (B, n) and returns (J, G):
    J a float64 device tensor of B values and G a (B, n) float64 device
    tensor (the batched ``FourDVar.evaluate`` / ``replay`` contract).
    """

    def __init__(self, evaluate, B, n, device, m=10, maxls=20):
        self.evaluate = evaluate
        self.B, self.n, self.m, self.maxls = int(B), int(n), int(m), int(maxls)
        self.dev = device
        dd = dict(dtype=torch.float64, device=device)
        # m + 1 slots per member: the newest pair is written into a free slot
        # and only then kept (or not, by the curvature test)
        self.S = torch.empty((self.B, self.m + 1, self.n), **dd)
        self.Y = torch.empty((self.B, self.m + 1, self.n), **dd)
        self.x = torch.empty((self.B, self.n), **dd)      # accepted iterate
        self.g = torch.empty((self.B, self.n), **dd)      # gradient there
        self.d = torch.empty((self.B, self.n), **dd)      # search direction
        self.xt = torch.empty((self.B, self.n), **dd)     # trial point
        self.q = torch.empty((self.B, self.n), **dd)      # two-loop work vector
        self.scal = torch.zeros(8 * self.B, **dd)         # device scalars of dot products
        self.scratch = torch.empty(_lib.VD_REDUCE_SCRATCH, **dd)

    # ------------------------------------------------------------ device ops
    def _s(self):
        return current_stream_ptr(self.dev)

    def _dots(self, pairs):
        """FP64 dot products ``[(a, b), ...]`` of device vectors, one host sync."""
        out = self.scal[:len(pairs)]
        out.zero_()
        for k, (a, b) in enumerate(pairs):
            _lib.call("vd_dot_f64", a.numel(), ptr(a), ptr(b), 1.0, ptr(out[k:k + 1]),
                      ptr(self.scratch), self._s())
        return out.cpu().numpy().tolist()

    def _axpby(self, a, x, b, y):
        _lib.call("vd_axpby", _lib.VD_F64, y.numel(), float(a), ptr(x), float(b), ptr(y),
                  self._s())

    def _direction(self, b, mem):
        """d_b = -H g_b by the two-loop recursion over member b's stored pairs."""
        q, d = self.q[b], self.d[b]
        q.copy_(self.g[b])
        k = len(mem.slots)
        if k == 0:
            self._axpby(0.0, q, -1.0, q)
            d.copy_(q)
            return
        alpha = [0.0] * k
        for i in reversed(range(k)):
            sl = mem.slots[i]
            alpha[i] = mem.rho[i] * self._dots([(self.S[b, sl], q)])[0]
            self._axpby(-alpha[i], self.Y[b, sl], 1.0, q)
        sl = mem.slots[-1]
        sy, yy = self._dots([(self.S[b, sl], self.Y[b, sl]), (self.Y[b, sl], self.Y[b, sl])])
        self._axpby(0.0, q, sy / yy, q)
        for i in range(k):
            sl = mem.slots[i]
            beta = mem.rho[i] * self._dots([(self.Y[b, sl], q)])[0]
            self._axpby(alpha[i] - beta, self.S[b, sl], 1.0, q)
        self._axpby(-1.0, q, 0.0, d)

    def _eval(self, X):
        J, G = self.evaluate(X)
        return J, G

    # ------------------------------------------------------------ driver
    def minimize(self, X0, maxiter=10, ftol=2.220446049250313e-09, gtol=1e-5, callback=None):
        """Minimise from ``X0`` ((B, n) or B*n values, host or device).  Returns
        ``(X, info)``: X the final iterates (device, (B, n)) and ``info`` a list
        of per-member dicts (cost history at accepted iterates, iterations,
        evaluations, stopping reason)."""
        B = self.B
        X0 = torch.as_tensor(X0, dtype=torch.float64, device=self.dev).reshape(B, self.n)
        self.x.copy_(X0)
        mems = [_Member(self.m) for _ in range(B)]
        J, G = self._eval(self.x)
        self.g.copy_(G.reshape(B, self.n))
        f = J.double().cpu().numpy().copy()
        nround = 1
        for b, mb in enumerate(mems):
            mb.nfev = 1
            mb.J_hist.append(float(f[b]))
            gmax = float(torch.max(torch.abs(self.g[b])))
            if gmax <= gtol:
                mb.done, mb.status = True, "CONVERGENCE: NORM OF PROJECTED GRADIENT <= PGTOL"
        # per-member line-search state
        fold = f.copy()
        gdold = np.zeros(B)
        dnorm = np.zeros(B)
        nls = np.zeros(B, dtype=int)

        def begin_iteration(b, mb):
            self._direction(b, mb)
            gd, dd_ = self._dots([(self.g[b], self.d[b]), (self.d[b], self.d[b])])
            if not gd < 0:
                # not a descent direction: drop the memory and take -g
                mb.slots, mb.rho = [], []
                self._direction(b, mb)
                gd, dd_ = self._dots([(self.g[b], self.d[b]), (self.d[b], self.d[b])])
                if not gd < 0:
                    mb.done, mb.status = True, "ABNORMAL"
                    return
            dnorm[b] = math.sqrt(dd_)
            gdold[b] = gd
            fold[b] = f[b]
            stp = min(1.0 / dnorm[b], mb.ls.stpmax) if mb.nit == 0 else 1.0
            mb.ls.start(f[b], gd, stp)
            nls[b] = 0
            self._trial(b, mb.ls.stp)

        for b, mb in enumerate(mems):
            if not mb.done:
                begin_iteration(b, mb)
        while not all(mb.done for mb in mems):
            for b, mb in enumerate(mems):
                if mb.done:
                    self.xt[b].copy_(self.x[b])
            J, G = self._eval(self.xt)
            nround += 1
            Jh = J.double().cpu().numpy()
            gd_t = self._dots([(G[b], self.d[b]) for b in range(B)])
            for b, mb in enumerate(mems):
                if mb.done:
                    continue
                mb.nfev += 1
                nls[b] += 1
                fb, gdb = float(Jh[b]), float(gd_t[b])
                if not (math.isfinite(fb) and math.isfinite(gdb)):
                    task = "NONFINITE"
                else:
                    task = mb.ls.next(fb, gdb)
                if task == "FG" and nls[b] < self.maxls:
                    self._trial(b, mb.ls.stp)
                    continue
                if task == "FG" or task == "NONFINITE":
                    # search failed: restart from the accepted point without memory
                    mb.fails += 1
                    if mb.fails >= 2 or not mb.slots:
                        mb.done, mb.status = True, "ABNORMAL"
                        continue
                    mb.slots, mb.rho = [], []
                    begin_iteration(b, mb)
                    continue
                # accepted: x = xt, g = G, update memory
                mb.fails = 0
                stp = mb.ls.stp
                sl = next(k for k in range(self.m + 1) if k not in mb.slots)
                self.S[b, sl].copy_(self.xt[b])
                self._axpby(-1.0, self.x[b], 1.0, self.S[b, sl])       # s = x_new - x
                self.Y[b, sl].copy_(G[b])
                self._axpby(-1.0, self.g[b], 1.0, self.Y[b, sl])       # y = g_new - g
                self.x[b].copy_(self.xt[b])
                self.g[b].copy_(G[b])
                f[b] = fb
                mb.nit += 1
                mb.J_hist.append(fb)
                if callback is not None:
                    callback(b, self.x[b])
                # scipy's driver tests the iteration count first, then L-BFGS-B
                # the projected gradient and the relative reduction
                if mb.nit >= maxiter:
                    mb.done, mb.status = True, "STOP: TOTAL NO. OF ITERATIONS REACHED LIMIT"
                    continue
                gmax = float(torch.max(torch.abs(self.g[b])))
                if gmax <= gtol:
                    mb.done, mb.status = True, "CONVERGENCE: NORM OF PROJECTED GRADIENT <= PGTOL"
                    continue
                if fold[b] - fb <= ftol * max(abs(fold[b]), abs(fb), 1.0):
                    mb.done, mb.status = True, "CONVERGENCE: RELATIVE REDUCTION OF F <= FACTR*EPSMCH"
                    continue
                # L-BFGS-B's curvature test: dr = s'y, ddum = -gdold * stp
                dr = (gdb - gdold[b]) * stp
                ddum = -gdold[b] * stp
                if dr > EPS * ddum:
                    sy = self._dots([(self.S[b, sl], self.Y[b, sl])])[0]
                    mb.slots.append(sl)
                    mb.rho.append(1.0 / sy)
                    if len(mb.slots) > self.m:
                        mb.slots.pop(0)
                        mb.rho.pop(0)
                begin_iteration(b, mb)
        info = [{"J": mb.J_hist, "nit": mb.nit, "nfev": mb.nfev, "status": mb.status}
                for mb in mems]
        return self.x, {"members": info, "rounds": nround}

    def _trial(self, b, stp):
        """xt_b = x_b + stp * d_b."""
        self.xt[b].copy_(self.x[b])
        self._axpby(float(stp), self.d[b], 1.0, self.xt[b])
