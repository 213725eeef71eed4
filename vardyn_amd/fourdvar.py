"""GPU-resident 4D-Var cost and gradient for the QG model, for one problem or a
batch of problems on the same grid.

Reference: ``Variational.cost_and_grad`` (mapping/src/tools_4Dvar.py:252-400)
driving ``Model_qg1l_jax.step/step_adj`` (mapping/src/mod.py:867-1076) over
the checkpoint schedule of ``Inv_4Dvar`` (mapping/src/inv.py:144-163),
datlas-ocean/VarDyn.

One evaluation follows the reference step by step:

forward, for each segment [c_i, c_{i+1}):
  misfit at t_i if it is an observation time (Jo += m.m / sigma_R^2),
  control -> flux where c_i % dtbasis == 0 (flux controls),
  segment start state kept (the trajectory checkpoint),
  QG step over the segment, then ``+ nstep*dt/86400 * Fssh``;
misfit at the final time; J = (Jo + Jb) / 2;
reverse, from the last segment to the first:
  ``adFssh += nstep*dt/86400 * adX1`` then the segment adjoint,
  basis transpose where c_i % dtbasis == 0, observation adjoint at t_i
  (fed with the raw misfit, as the reference does, tools_4Dvar.py:349,380).

Controls: ``"ic"`` -- the control vector is the initial SSH (the reference
computes this adjoint and drops it); ``"offset"`` -- the reference's
``Basis_offset`` flux (basis.py:3939-4000); ``"basis"`` -- the SSH flux of a
time-dependent reduced basis (``basis=``, e.g. the balanced-motion wavelets of
``vardyn_amd.basis_bm.Basis_bm``) with per-coefficient prior standard
deviations ``sigma_b`` (its Q): the fields of all basis times come from one
batched synthesis before the forward sweep, and the adjoint fields of all
basis times go through one batched transpose after the reverse sweep.

Tape: ``"full"`` keeps every stage state of the window in HBM during the
forward sweep (no recomputation); ``"checkpoint"`` keeps only the segment
start states and recomputes each segment into one reusable tape during the
reverse sweep, like the reference's ``jax.vjp`` of a segment; ``"auto"``
takes the full tape when it fits in 80% of the free device memory.  All give
identical results.

Batches: a plan built with ``nbatch = B`` advances B independent problems in
the same kernel launches.  Each member has its own observation operator,
control vector, J and G; the members share the grid, the time step and the
checkpoint schedule (their observations fall on the same model steps, as for
regions seen by the same satellite passes).

All arithmetic is on the device: cost accumulators and the control/gradient
vectors are float64; model fields use the plan precision.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._lib import VardynError
from .qg import QgPlan, current_stream_ptr, ptr


class FourDVar:
    def __init__(self, plan: QgPlan, obsop, checkpoints, T, control="ic", sigma_b=None,
                 Xb=None, sigma_r=None, prec=False, dtbasis=None, tape="checkpoint", hb=None,
                 land=None, basis=None):
        if control not in ("ic", "offset", "basis"):
            raise ValueError("control must be 'ic', 'offset' or 'basis'")
        if (control == "basis") != (basis is not None):
            raise ValueError("control 'basis' takes a basis (and only it does)")
        if tape not in ("checkpoint", "full", "auto"):
            raise ValueError("tape must be 'checkpoint', 'full' or 'auto'")
        if sigma_b is not None and not np.isscalar(sigma_b) and control != "basis":
            raise NotImplementedError("a per-coefficient sigma_b needs control 'basis'")
        if prec and sigma_b is None:
            raise ValueError("preconditioning needs sigma_b")
        self.plan = plan
        self.B = plan.nb
        self.Hs = list(obsop) if isinstance(obsop, (list, tuple)) else [obsop]
        if len(self.Hs) != self.B:
            raise VardynError(f"{len(self.Hs)} observation operators for {self.B} batch members")
        self.H = self.Hs[0]
        self.dev = plan.device
        self.dtype = plan.dtype
        self.control = control
        self.sigma_b = sigma_b
        self.sigma_r = 1.0 if sigma_r is None else float(sigma_r)
        self.prec = bool(prec)
        cps = np.asarray(checkpoints, dtype=np.int64)
        if cps[0] != 0 or np.any(np.diff(cps) <= 0):
            raise VardynError("checkpoints must start at 0 and increase")
        self.cps = cps
        self.T = np.asarray(T, dtype=np.float64)
        if len(self.T) != cps[-1] + 1:
            raise VardynError("T must hold one time per model step (len == checkpoints[-1] + 1)")
        self.dtbasis = int(dtbasis) if dtbasis is not None else int(np.max(np.diff(cps)))
        self.nseg = len(cps) - 1
        self.nsteps = np.diff(cps).astype(int)
        self.n = plan.ny * plan.nx
        self.nctrl = self.n if control == "ic" else (basis.nbasis if control == "basis" else 1)
        if tape == "auto":
            # the whole-window tape when it takes at most 80% of the free HBM
            need = sum(plan.tape_bytes(int(k)) for k in self.nsteps)
            free, _ = torch.cuda.mem_get_info(self.dev)
            tape = "full" if need <= 0.8 * free else "checkpoint"
        self.tape_mode = tape

        # observation table index of every member at each segment start / final time
        def obs_index(H, t):
            return H.index(t) if H.is_obs_time(t) else -1
        self.seg_obs = [[obs_index(H, self.T[c]) for H in self.Hs] for c in cps[:-1]]
        self.final_obs = [obs_index(H, self.T[cps[-1]]) for H in self.Hs]
        for row in self.seg_obs + [self.final_obs]:
            if len({r >= 0 for r in row}) > 1:
                raise VardynError("batch members must observe on the same model steps")

        fshape = (plan.ny, plan.nx) if self.B == 1 else (self.B, plan.ny, plan.nx)
        dd = dict(device=self.dev)
        self.fshape = fshape
        # None: zero boundary field (closed basin, psi = 0); the plan then skips h_b, q_b
        self.hb = None if hb is None else plan.field(hb)
        self.states = torch.empty((self.nseg + 1,) + fshape, dtype=self.dtype, **dd)
        if tape == "full":
            es = plan.empty().element_size()
            self.tape_elems = [plan.tape_bytes(int(k)) // es for k in self.nsteps]
            self.tape_off = np.concatenate([[0], np.cumsum(self.tape_elems)]).astype(np.int64)
            self.tape = torch.empty(int(self.tape_off[-1]), dtype=self.dtype, **dd)
        else:
            self.tape = plan.new_tape(int(self.nsteps.max()))
        self.lam = torch.empty(fshape, dtype=self.dtype, **dd)
        self.lam2 = torch.empty(fshape, dtype=self.dtype, **dd)
        self.params = torch.zeros(fshape, dtype=self.dtype, **dd)
        self.adparams = torch.zeros(fshape, dtype=self.dtype, **dd)
        self.ones64 = torch.ones(self.n, dtype=torch.float64, **dd)
        self.p64 = torch.empty((self.B, self.n), dtype=torch.float64, **dd)
        self.X = torch.empty((self.B, self.nctrl), dtype=torch.float64, **dd)
        self.G = torch.empty((self.B, self.nctrl), dtype=torch.float64, **dd)
        self.J = torch.zeros(self.B, dtype=torch.float64, **dd)
        self.scratch = torch.empty(_lib.VD_REDUCE_SCRATCH, dtype=torch.float64, **dd)
        if Xb is None:
            self.Xb = torch.zeros((self.B, self.nctrl), dtype=torch.float64, **dd)
        else:
            self.Xb = torch.as_tensor(np.asarray(Xb, dtype=np.float64), device=self.dev) \
                .reshape(self.B, self.nctrl).clone()
        self.var0 = torch.zeros(fshape, dtype=self.dtype, **dd)
        self.land = land
        self._graph = None
        self.basis = basis
        self.sig_vec = None
        if control == "basis":
            if sigma_b is None:
                raise ValueError("control 'basis' needs sigma_b (the basis's Q)")
            sb = np.broadcast_to(np.asarray(sigma_b, dtype=np.float64), (self.nctrl,))
            self.sig_vec = torch.as_tensor(np.ascontiguousarray(sb), **dd)
            self.inv_s2_vec = 1.0 / self.sig_vec ** 2
            self.wtmp = torch.empty((self.B, self.nctrl), dtype=torch.float64, **dd)
            # segments that start at a basis time, and their times in days
            self.bseg = [i for i in range(self.nseg) if self.cps[i] % self.dtbasis == 0]
            self.bindex = {i: k for k, i in enumerate(self.bseg)}
            self.bprep = basis.prepare(self.T[self.cps[self.bseg]] / (3600 * 24))
            nb = len(self.bseg)
            self.bfields = torch.empty((self.B, nb, plan.ny, plan.nx), dtype=torch.float64, **dd)
            self.badj = torch.zeros((self.B, nb, plan.ny, plan.nx), dtype=torch.float64, **dd)

    # ------------------------------------------------------------ helpers
    def _s(self):
        return current_stream_ptr(self.dev)

    def _tape(self, i):
        if self.tape_mode == "full":
            return self.tape[int(self.tape_off[i]):int(self.tape_off[i + 1])]
        return self.tape

    def _axpby(self, a, x, b, y):
        prec = _lib.VD_F64 if y.dtype == torch.float64 else _lib.VD_F32
        _lib.call("vd_axpby", prec, y.numel(), float(a), ptr(x), float(b), ptr(y), self._s())

    def _dot(self, x, y, alpha, out):
        _lib.call("vd_dot_f64", x.numel(), ptr(x), ptr(y), float(alpha), ptr(out),
                  ptr(self.scratch), self._s())

    def _convert(self, src, dst, nan_to_zero=False):
        sp = _lib.VD_F64 if src.dtype == torch.float64 else _lib.VD_F32
        dp = _lib.VD_F64 if dst.dtype == torch.float64 else _lib.VD_F32
        _lib.call("vd_convert", sp, ptr(src), dp, ptr(dst), src.numel(), 1 if nan_to_zero else 0,
                  self._s())

    def _flux_scale(self, i):
        return self.nsteps[i] * self.plan.dt / (3600 * 24)

    def _member(self, field, b):
        return field if self.B == 1 else field[b]

    def _misfits(self, idx, state, J, inv_r2):
        for b, H in enumerate(self.Hs):
            H.misfit_dev(idx[b], self._member(state, b), jo=J[b:b + 1], inv_sigma_r2=inv_r2)

    def _obs_adj(self, idx, lam):
        for b, H in enumerate(self.Hs):
            H.adj_dev(idx[b], H.tables[idx[b]].misfit, self._member(lam, b))

    # --------------------------------------------------------- evaluation
    def run(self, X0):
        """Enqueue one cost+gradient evaluation of every member.

        ``X0``: float64 device tensor with ``B * nctrl`` elements.  Results are
        left in ``self.J`` (B values) and ``self.G`` (B x nctrl).
        """
        plan = self.plan
        if X0.dtype != torch.float64 or X0.device != self.dev or X0.numel() != self.B * self.nctrl:
            raise VardynError(f"control must be a float64 device tensor of "
                              f"{self.B * self.nctrl} elements")
        X0 = X0.reshape(self.B, self.nctrl).contiguous()
        J, X = self.J, self.X
        J.zero_()
        inv_sb2 = 0.0 if (self.sigma_b is None or self.sig_vec is not None) \
            else 1.0 / float(self.sigma_b) ** 2
        # background term (tools_4Dvar.py:262-271); J accumulates (Jo + Jb) / 2
        self._convert(self.Xb, X)
        if self.sig_vec is not None:
            # per-coefficient standard deviations (tools_4Dvar.py:262-271)
            if self.prec:
                X.addcmul_(self.sig_vec, X0)
                for b in range(self.B):
                    self._dot(X0[b], X0[b], 0.5, J[b:b + 1])
            else:
                X.add_(X0)
                torch.mul(X0, self.inv_s2_vec, out=self.wtmp)
                for b in range(self.B):
                    self._dot(X0[b], self.wtmp[b], 0.5, J[b:b + 1])
        elif self.sigma_b is not None:
            self._axpby(self.sigma_b if self.prec else 1.0, X0, 1.0, X)
            for b in range(self.B):
                self._dot(X0[b], X0[b], 0.5 if self.prec else 0.5 * inv_sb2, J[b:b + 1])
        else:
            self._axpby(1.0, X0, -1.0, X)
        inv_r2 = 0.5 / self.sigma_r ** 2
        s0 = self.states[0]
        if self.control == "ic":
            self._convert(X, s0)
            if self.land is not None:
                _lib.call("vd_qg_fill_land", plan._h, ptr(s0), float("nan"), self._s())
        else:
            s0.copy_(self.var0)
        params = self.params
        params.zero_()
        if self.control == "basis":
            for b in range(self.B):
                self.basis.operg_prepared(self.bprep, X[b], self.bfields[b])
        # forward sweep
        for i in range(self.nseg):
            st, nxt = self.states[i], self.states[i + 1]
            if self.seg_obs[i][0] >= 0:
                self._misfits(self.seg_obs[i], st, J, inv_r2)
            if self.control == "offset" and self.cps[i] % self.dtbasis == 0:
                # Basis_offset.operg: params = X everywhere (basis.py:3968-3984)
                params.copy_(X[:, 0].to(params.dtype).reshape(self.B, 1, 1)
                             .expand(self.B, plan.ny, plan.nx).reshape(self.fshape))
            if self.control == "basis" and i in self.bindex:
                # basis.operg(t, X, State=params) (tools_4Dvar.py:295-299)
                params.copy_(self.bfields[:, self.bindex[i]].reshape(self.fshape))
            tape = self._tape(i) if self.tape_mode == "full" else None
            plan.step(st, self.hb, int(self.nsteps[i]), h1=nxt, tape=tape)
            if self.control in ("offset", "basis"):
                self._axpby(self._flux_scale(i), params, 1.0, nxt)
        if self.final_obs[0] >= 0:
            self._misfits(self.final_obs, self.states[-1], J, inv_r2)

        # reverse sweep
        lam, lam2 = self.lam, self.lam2
        lam.zero_()
        adpar = self.adparams
        adpar.zero_()
        G = self.G
        G.zero_()
        if self.final_obs[0] >= 0:
            self._obs_adj(self.final_obs, lam)
        for i in reversed(range(self.nseg)):
            if self.control in ("offset", "basis"):
                self._axpby(self._flux_scale(i), lam, 1.0, adpar)
            plan.step_adj(lam, self.states[i], self.hb, int(self.nsteps[i]), adh0=lam2,
                          tape=self._tape(i), tape_valid=self.tape_mode == "full")
            lam, lam2 = lam2, lam
            if self.control == "offset" and self.cps[i] % self.dtbasis == 0:
                self._convert(adpar, self.p64)
                for b in range(self.B):
                    self._dot(self.p64[b], self.ones64, 1.0, G[b])
                adpar.zero_()
            if self.control == "basis" and i in self.bindex:
                # adX += basis.operg_transpose(t, adState.params) (tools_4Dvar.py:369-374),
                # gathered here and transposed for all basis times after the sweep
                self.badj[:, self.bindex[i]].copy_(adpar.reshape(self.B, plan.ny, plan.nx))
                adpar.zero_()
            if self.seg_obs[i][0] >= 0:
                self._obs_adj(self.seg_obs[i], lam)
        # gradient (tools_4Dvar.py:388-398)
        if self.control == "ic":
            self._convert(lam, G, nan_to_zero=True)
        if self.control == "basis":
            for b in range(self.B):
                self.basis.operg_transpose_prepared(self.bprep, self.badj[b], G[b])
            if self.prec:
                G.mul_(self.sig_vec)
                G.add_(X0)
            else:
                G.addcmul_(X0, self.inv_s2_vec)
            return J, G
        if self.prec:
            self._axpby(0.0, G, float(self.sigma_b), G)
            self._axpby(1.0, X0, 1.0, G)
        elif self.sigma_b is not None:
            self._axpby(inv_sb2, X0, 1.0, G)
        return J, G

    # ------------------------------------------------------------ graphs
    def capture(self):
        """Capture one whole evaluation in a HIP graph.

        Every buffer the evaluation touches is preallocated, so the kernel
        sequence is static; replaying the graph replaces tens of thousands of
        individual launches per gradient.  An eager evaluation runs first so
        that lazily created buffers exist before capture.
        """
        self._x_static = torch.zeros(self.B * self.nctrl, dtype=torch.float64, device=self.dev)
        self.run(self._x_static)
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run(self._x_static)
        torch.cuda.synchronize(self.dev)
        self._graph = g
        return self

    @property
    def captured(self):
        return self._graph is not None

    def replay(self, X0):
        """Run the captured evaluation on control ``X0`` (float64 device tensor)."""
        if not self.captured:
            raise VardynError("capture() first")
        self._x_static.copy_(X0.reshape(-1))
        self._graph.replay()
        return self.J, self.G

    def evaluate(self, X0):
        """Graph replay when captured, eager launches otherwise."""
        return self.replay(X0) if self.captured else self.run(X0)

    def cost_and_grad(self, X0):
        """Host interface for scipy: numpy in, (J, G) out.

        For one problem J is a float and G has ``nctrl`` values; for a batch
        J has B values and G is (B, nctrl).
        """
        x = torch.as_tensor(np.asarray(X0, dtype=np.float64), device=self.dev)
        J, G = self.evaluate(x)
        if self.B == 1:
            return float(J[0].item()), G[0].cpu().numpy()
        return J.cpu().numpy(), G.cpu().numpy()
