"""Uneven problem sets in shared launches: 4D-Var cost and gradient of many
independent QG problems with their own grid sizes, gravity-wave speeds,
windows, observations and checkpoint schedules.

The problems are ordered by window length (longest first) and advanced step
by step.  At global step s the members still running form a prefix, and each
RK stage, Helmholtz inversion and adjoint stage of that step is one launch of
the ragged kernels over the prefix (``vd_group_*``): the launch count of an
evaluation is that of the longest problem, not the sum over problems.  What
differs between members -- segment starts and ends of their own checkpoint
schedules, observation times -- runs as small per-member launches
(``vd_qg_prep/finish/adj_init/adj_final``, the observation operator) at the
steps where it happens.

Per member, the sequence of operations is exactly that of ``FourDVar`` with a
full tape (``Variational.cost_and_grad``, mapping/src/tools_4Dvar.py:252-400,
over the segments of mapping/src/inv.py:144-163, datlas-ocean/VarDyn), and
every kernel computes a member exactly as its own plan does with GEMM
configuration 0, so the results equal the single-problem engine's bit for
bit.  Scope: initial-state control, zero boundary fields, full trajectory tape
(sized per step to the members still running).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import VardynError
from .qg import QgPlan, current_stream_ptr, ptr


class QgGroup:
    """Plans of the members (single-problem ``QgPlan``s) joined into a ragged group."""

    def __init__(self, plans):
        self.plans = list(plans)
        if not self.plans:
            raise VardynError("empty group")
        arr = (C.c_void_p * len(self.plans))(*[p._h for p in self.plans])
        h = C.c_void_p()
        _lib.call("vd_group_create", arr, len(self.plans), C.byref(h))
        self._h = h
        offs = (C.c_size_t * (len(self.plans) + 1))()
        _lib.call("vd_group_offsets", self._h, offs)
        self.offsets = [int(v) for v in offs]
        self.dtype = self.plans[0].dtype
        self.device = self.plans[0].device

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            try:
                _lib.load().vd_group_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def n(self):
        return len(self.plans)

    def _s(self):
        return current_stream_ptr(self.device)

    def stage(self, nact, h, q, qbase, k0, k1, c0, c1, c2, qout, kstore=None):
        _lib.call("vd_group_stage", self._h, int(nact), h, q, qbase, k0, k1, float(c0), float(c1),
                  float(c2), qout, kstore, self._s())

    def inversion(self, nact, F, out, accumulate=False, Fsub=None, hb=None):
        _lib.call("vd_group_inversion", self._h, int(nact), F, Fsub, hb, out,
                  1 if accumulate else 0, self._s())

    def stage_adj(self, nact, h, q, s0, s1, s2, a0, a1, a2, acc0, acc1, acc2, lq_out, lh_out):
        _lib.call("vd_group_stage_adj", self._h, int(nact), h, q, s0, s1, s2, float(a0),
                  float(a1), float(a2), acc0, acc1, acc2, lq_out, lh_out, self._s())

    def prep(self, nact, h0, hb, h, q, hbm=None, qb=None):
        _lib.call("vd_group_prep", self._h, int(nact), h0, hb, h, q, hbm, qb, self._s())

    def finish(self, nact, h, h1):
        _lib.call("vd_group_finish", self._h, int(nact), h, h1, self._s())

    def adj_init(self, nact, adh1, lh, lq):
        _lib.call("vd_group_adj_init", self._h, int(nact), adh1, lh, lq, self._s())

    def adj_final(self, nact, hcur, lh, lq, adh0):
        _lib.call("vd_group_adj_final", self._h, int(nact), hcur, lh, lq, adh0, self._s())

    def autotune(self, reps=5):
        """Pick the group's tile shapes and inversion form by timing them on its
        members (``vd_group_autotune``); returns the choices."""
        ch = (C.c_int * 6)()
        _lib.call("vd_group_autotune", self._h, int(reps), ch, self._s())
        return {"split_tiles": [("64x32", "32x32")[c] if c >= 0 else None for c in ch[:4]],
                "adjoint_tile": ("64x16", "32x32")[ch[4]],
                "fused_small_inversion": None if ch[5] < 0 else bool(ch[5])}

    def set_tuning(self, tuning):
        """Apply choices in the form ``autotune`` returns."""
        ch = (C.c_int * 6)(*[-1 if t is None else ("64x32", "32x32").index(t) for t in tuning["split_tiles"]],
                           ("64x16", "32x32").index(tuning["adjoint_tile"]),
                           -1 if tuning["fused_small_inversion"] is None else int(tuning["fused_small_inversion"]))
        _lib.call("vd_group_set_tuning", self._h, ch)


class RaggedFourDVar:
    """Cost and gradient of every problem of ``problems`` in one step-synchronous sweep.

    ``problems``: objects with ``plan`` (a single-problem ``QgPlan``), ``H``
    (``Obsop_interp_l3``), ``checkpoints``, ``T`` and ``case.sigma_b`` -- what
    ``vardyn_amd.problem.build_problem`` returns.
    """

    def __init__(self, problems, autotune=False):
        self.problems = list(problems)
        n = len(self.problems)
        if n == 0:
            raise VardynError("no problems")
        p0 = self.problems[0].plan
        for pb in self.problems:
            if pb.plan.nb != 1:
                raise VardynError("members must be single-problem plans")
            if pb.plan.dtype != p0.dtype or pb.plan.dt != p0.dt:
                raise VardynError("members need the same precision and time step")
        self.dev, self.dtype = p0.device, p0.dtype
        self.dt = float(p0.dt)
        self.scheme = p0.time_scheme
        # members ordered by window length, longest first (stable)
        nsteps = [int(pb.checkpoints[-1]) for pb in self.problems]
        self.order = sorted(range(n), key=lambda i: -nsteps[i])
        self.pos = {i: j for j, i in enumerate(self.order)}
        self.members = [self.problems[i] for i in self.order]
        self.nst = np.array([nsteps[i] for i in self.order])
        self.S = int(self.nst.max())
        self.group = QgGroup([m.plan for m in self.members])
        # tile shapes / inversion form timed on these members (None: defaults)
        self.tuning = self.group.autotune() if autotune else None
        off = self.group.offsets
        self.nf = [off[j + 1] - off[j] for j in range(n)]
        es = torch.empty((), dtype=self.dtype).element_size()
        self.es = es
        self.nstage = {"Euler": 1, "rk2": 2, "rk3": 3}[self.scheme]
        # active prefixes: K0(s) members hold a state at step s, K1(s) take step s
        self.K0 = [int(np.sum(self.nst >= s)) for s in range(self.S + 1)]
        self.K1 = [int(np.sum(self.nst > s)) for s in range(self.S)]
        # per member segment events and observation indices (as FourDVar)
        self.starts = [[] for _ in range(self.S + 1)]
        self.ends = [[] for _ in range(self.S + 1)]
        self.seg_obs, self.final_obs = [], []
        for j, m in enumerate(self.members):
            cps = np.asarray(m.checkpoints, dtype=np.int64)
            T = np.asarray(m.T, dtype=np.float64)
            for i in range(len(cps) - 1):
                self.starts[int(cps[i])].append((j, i))
                self.ends[int(cps[i + 1])].append((j, i + 1))
            self.seg_obs.append([m.H.index(T[c]) if m.H.is_obs_time(T[c]) else -1
                                 for c in cps[:-1]])
            self.final_obs.append(m.H.index(T[cps[-1]]) if m.H.is_obs_time(T[cps[-1]]) else -1)
        # trajectory tape: per step, stage-0 state of the K0(s) members, then the
        # further stage states of the K1(s) members; h and q of each
        F = lambda k: off[k]  # noqa: E731  (elements of the first k members)
        self.slot = {}
        pos = 0
        for s in range(self.S + 1):
            self.slot[(s, 0)] = pos
            pos += 2 * F(self.K0[s])
            if s < self.S:
                for st in range(1, self.nstage):
                    self.slot[(s, st)] = pos
                    pos += 2 * F(self.K1[s])
        dd = dict(dtype=self.dtype, device=self.dev)
        self.tape = torch.empty(max(pos, 1), **dd)
        self.tape_bytes = pos * es
        ntot = off[-1]
        self.scr = {k: torch.zeros(max(ntot, 1), **dd)
                    for k in ("lh", "lq", "lqx", "lq1", "lq2", "lhs", "k0", "k1")}
        self.states = [torch.empty((len(m.checkpoints), m.plan.ny, m.plan.nx), **dd)
                       for m in self.members]
        self.lam = [torch.zeros((m.plan.ny, m.plan.nx), **dd) for m in self.members]
        self.lam2 = [torch.zeros((m.plan.ny, m.plan.nx), **dd) for m in self.members]
        # controls and gradients in the caller's problem order
        self.nctrl = [pb.plan.ny * pb.plan.nx for pb in self.problems]
        self.coff = np.concatenate([[0], np.cumsum(self.nctrl)]).astype(np.int64)
        self.X = torch.zeros(int(self.coff[-1]), dtype=torch.float64, device=self.dev)
        self.G = torch.zeros(int(self.coff[-1]), dtype=torch.float64, device=self.dev)
        self.J = torch.zeros(n, dtype=torch.float64, device=self.dev)
        self.scratch = torch.empty(_lib.VD_REDUCE_SCRATCH, dtype=torch.float64, device=self.dev)
        self.sigma_b = [float(pb.case.sigma_b) for pb in self.problems]
        self._graph = None

    # ------------------------------------------------------------ helpers
    def _s(self):
        return current_stream_ptr(self.dev)

    def _h(self, s, st):
        return self.tape.data_ptr() + self.slot[(s, st)] * self.es

    def _q(self, s, st):
        k = self.K0[s] if st == 0 else self.K1[s]
        return self._h(s, st) + self.group.offsets[k] * self.es

    def _view(self, base_ptr, j):
        """Address of member j inside a flat field that starts at base_ptr."""
        return base_ptr + self.group.offsets[j] * self.es

    def _sp(self, name):
        return self.scr[name].data_ptr()

    def _ctrl(self, t, i):
        return t[int(self.coff[i]):int(self.coff[i + 1])]

    def _member_tensor(self, base_ptr, j):
        return base_ptr + self.group.offsets[j] * self.es

    # --------------------------------------------------------- evaluation
    def run(self, X0):
        """Enqueue one evaluation of every problem; ``X0`` is a float64 device
        tensor with the problems' controls concatenated in their given order.
        Results: ``self.J`` (one value per problem) and ``self.G`` (same layout
        as ``X0``)."""
        if X0.dtype != torch.float64 or X0.device != self.dev or X0.numel() != self.X.numel():
            raise VardynError(f"control must be a float64 device tensor of {self.X.numel()} "
                              f"elements")
        g, dt, s_ = self.group, self.dt, self._s()
        J = self.J
        J.zero_()
        # background term and initial states (x_b = 0, tools_4Dvar.py:262-271)
        for i, pb in enumerate(self.problems):
            x = self._ctrl(X0, i)
            inv_sb2 = 1.0 / self.sigma_b[i] ** 2
            _lib.call("vd_dot_f64", x.numel(), ptr(x), ptr(x), 0.5 * inv_sb2,
                      J[i:i + 1].data_ptr(), ptr(self.scratch), s_)
            j = self.pos[i]
            st0 = self.states[j][0]
            _lib.call("vd_convert", _lib.VD_F64, ptr(x), self._prec(), ptr(st0), x.numel(), 0, s_)
        # forward sweep
        for s in range(self.S):
            for j, i in self.starts[s]:
                m, st = self.members[j], self.states[j][i]
                if self.seg_obs[j][i] >= 0:
                    self._misfit(j, self.seg_obs[j][i], st)
                _lib.call("vd_qg_prep", m.plan._h, ptr(st), None, self._view(self._h(s, 0), j),
                          self._view(self._q(s, 0), j), s_)
            k = self.K1[s]
            hc, qc = self._h(s, 0), self._q(s, 0)
            hn, qn = self._h(s + 1, 0), self._q(s + 1, 0)
            if self.scheme == "Euler":
                g.stage(k, hc, qc, qc, None, None, 0, 0, dt, qn)
            elif self.scheme == "rk2":
                hs1, qs1 = self._h(s, 1), self._q(s, 1)
                g.stage(k, hc, qc, qc, None, None, 0, 0, 0.5 * dt, qs1)
                g.inversion(k, qs1, hs1)
                g.stage(k, hs1, qs1, qc, None, None, 0, 0, dt, qn)
            else:
                hs1, qs1, hs2, qs2 = self._h(s, 1), self._q(s, 1), self._h(s, 2), self._q(s, 2)
                k0, k1 = self._sp("k0"), self._sp("k1")
                g.stage(k, hc, qc, qc, None, None, 0, 0, 0.5 * dt, qs1, k0)
                g.inversion(k, qs1, hs1)
                g.stage(k, hs1, qs1, qc, k0, None, -dt, 0, 2 * dt, qs2, k1)
                g.inversion(k, qs2, hs2)
                g.stage(k, hs2, qs2, qc, k0, k1, dt / 6, 4 * dt / 6, dt / 6, qn)
            g.inversion(k, qn, hn)
            for j, i in self.ends[s + 1]:
                _lib.call("vd_qg_finish", self.members[j].plan._h,
                          self._view(self._h(s + 1, 0), j), ptr(self.states[j][i]), s_)
        for j in range(len(self.members)):
            if self.final_obs[j] >= 0:
                self._misfit(j, self.final_obs[j], self.states[j][-1])

        # reverse sweep
        lam, lam2 = self.lam, self.lam2
        for j in range(len(self.members)):
            lam[j].zero_()
            if self.final_obs[j] >= 0:
                self._obs_adj(j, self.final_obs[j], lam[j])
        lh, lq, lqx = self._sp("lh"), self._sp("lq"), self._sp("lqx")
        lq1, lq2, lhs = self._sp("lq1"), self._sp("lq2"), self._sp("lhs")
        for s in reversed(range(self.S)):
            for j, i in self.ends[s + 1]:
                _lib.call("vd_qg_adj_init", self.members[j].plan._h, ptr(lam[j]),
                          self._view(lh, j), self._view(lq, j), s_)
            k = self.K1[s]
            g.inversion(k, lh, lq, accumulate=True)
            h0s, q0s = self._h(s, 0), self._q(s, 0)
            if self.scheme == "Euler":
                g.stage_adj(k, h0s, q0s, lq, None, None, dt, 0, 0, lq, None, None, lqx, lh)
                lq, lqx = lqx, lq
            elif self.scheme == "rk2":
                g.stage_adj(k, self._h(s, 1), self._q(s, 1), lq, None, None, dt, 0, 0,
                            None, None, None, lq1, lhs)
                g.inversion(k, lhs, lq1, accumulate=True)
                g.stage_adj(k, h0s, q0s, lq1, None, None, 0.5 * dt, 0, 0, lq, lq1, None, lq, lh)
            else:
                g.stage_adj(k, self._h(s, 2), self._q(s, 2), lq, None, None, dt / 6, 0, 0,
                            None, None, None, lq2, lhs)
                g.inversion(k, lhs, lq2, accumulate=True)
                g.stage_adj(k, self._h(s, 1), self._q(s, 1), lq, lq2, None, 4 * dt / 6, 2 * dt,
                            0, None, None, None, lq1, lhs)
                g.inversion(k, lhs, lq1, accumulate=True)
                g.stage_adj(k, h0s, q0s, lq, lq2, lq1, dt / 6, -dt, 0.5 * dt, lq, lq2, lq1,
                            lqx, lh)
                lq, lqx = lqx, lq
            for j, i in self.starts[s]:
                _lib.call("vd_qg_adj_final", self.members[j].plan._h, self._view(h0s, j),
                          self._view(lh, j), self._view(lq, j), ptr(lam2[j]), s_)
                lam[j], lam2[j] = lam2[j], lam[j]
                if self.seg_obs[j][i] >= 0:
                    self._obs_adj(j, self.seg_obs[j][i], lam[j])
        # gradients (tools_4Dvar.py:388-398): G = adX + X0 / sigma_b^2
        for i in range(len(self.problems)):
            j = self.pos[i]
            gi = self._ctrl(self.G, i)
            _lib.call("vd_convert", self._prec(), ptr(lam[j]), _lib.VD_F64, ptr(gi), gi.numel(),
                      1, s_)
            _lib.call("vd_axpby", _lib.VD_F64, gi.numel(), 1.0 / self.sigma_b[i] ** 2,
                      ptr(self._ctrl(X0, i)), 1.0, ptr(gi), s_)
        return self.J, self.G

    def _prec(self):
        return _lib.VD_F64 if self.dtype == torch.float64 else _lib.VD_F32

    def _misfit(self, j, idx, state):
        i = self.order[j]
        self.members[j].H.misfit_dev(idx, state, jo=self.J[i:i + 1], inv_sigma_r2=0.5)

    def _obs_adj(self, j, idx, lam):
        H = self.members[j].H
        H.adj_dev(idx, H.tables[idx].misfit, lam)

    # ------------------------------------------------------------ graphs
    def capture(self):
        """Capture one whole evaluation of every problem in a HIP graph."""
        self._x_static = torch.zeros_like(self.X)
        self.run(self._x_static)
        torch.cuda.synchronize(self.dev)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self.run(self._x_static)
        torch.cuda.synchronize(self.dev)
        self._graph = graph
        return self

    def evaluate(self, X0):
        if self._graph is None:
            return self.run(X0)
        self._x_static.copy_(X0.reshape(-1))
        self._graph.replay()
        return self.J, self.G

    def split(self, flat):
        """Per-problem views of a flat control/gradient tensor."""
        return [self._ctrl(flat, i) for i in range(len(self.problems))]
