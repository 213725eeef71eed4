"""GPU-resident 4D-Var cost and gradient for the multilayer shallow-water model
with a time-dependent reduced basis as the control.

Reference: ``Variational.cost_and_grad`` (mapping/src/tools_4Dvar.py:252-400)
driving ``Model_qgsw`` (mapping/src/mod.py:2412-3617) with ``nl > 1``,
datlas-ocean/VarDyn:

* the model state is the per-layer (u, v, h) of ``SW`` (sw.py:67-937), the
  observed surface field is SSH = sum of the layer thicknesses
  (``_diagnose_surface``, mod.py:3012-3016);
* the control is a reduced basis (e.g. ``Basis_wavelet3d``) on the SSH
  parameter: at each basis time the basis field F (SSH units per day)
  becomes the model's thickness forcing of the top layer, Fh = F / 86400
  (``jstep_core``, mod.py:3144-3210), added as h += dt Fh after every step
  (sw.py:875-877) until the next basis time;
* the reverse sweep follows ``step_adj`` (mod.py:3455-3617): the surface
  adjoint (the observation adjoints) goes into every layer's thickness
  adjoint at the start of each segment's adjoint and is then cleared; the
  forcing cotangent of the top layer, / 86400, accumulates into the SSH
  parameter adjoint, which the basis transpose consumes at basis times
  (tools_4Dvar.py:369-374).

The forward sweep keeps the trajectory (``vd_sw_step_keep``): every step's
state and RK3 stage states with ``tape="stages"``, every step's state with
``tape="full"``, or every ceil(sqrt(n))-th state of each segment with
``tape="checkpoint"``; the segment adjoint (``vd_sw_step_adj_kept``) runs
the hand-written adjoint RK3 stages from it, recomputing what was not kept.  The fields of all basis times come from one batched synthesis
before the forward sweep; the adjoint fields of all basis times go through
one batched transpose after the reverse sweep.  Cost and gradient are FP64;
the model runs in FP64.

Layout: SW arrays are x before y (u (nl, nx+1, ny), v (nl, nx, ny+1),
h (nl, nx, ny)), so the observation operator and the basis are built on the
(nx, ny) h grid (the State grid transposed): their flat indices address the
model's arrays directly.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._lib import VardynError
from .qg import current_stream_ptr, ptr
from .sw import SW


class SwFourDVar:
    def __init__(self, model: SW, obsop, checkpoints, T, basis, x0, sigma_b, Xb=None, sigma_r=None,
                 prec=False, dtbasis=None, tape="auto"):
        self.model = model
        self.H = obsop
        self.basis = basis
        self.dev = model.device
        cps = np.asarray(checkpoints, dtype=np.int64)
        if cps[0] != 0 or np.any(np.diff(cps) <= 0):
            raise VardynError("checkpoints must start at 0 and increase")
        self.cps = cps
        self.T = np.asarray(T, dtype=np.float64)
        if len(self.T) != cps[-1] + 1:
            raise VardynError("T must hold one time per model step (len == checkpoints[-1] + 1)")
        self.nseg = len(cps) - 1
        self.nsteps = np.diff(cps).astype(int)
        self.dtbasis = int(dtbasis) if dtbasis is not None else int(np.max(self.nsteps))
        if tuple(basis.shape_phys) != (model.nx, model.ny):
            raise VardynError("the basis must live on the model's (nx, ny) h grid")
        self.nctrl = int(basis.nbasis)
        self.prec = bool(prec)
        self.sigma_r = 1.0 if sigma_r is None else float(sigma_r)
        dd = dict(device=self.dev)
        sb = np.broadcast_to(np.asarray(sigma_b, dtype=np.float64), (self.nctrl,))
        self.sig_vec = torch.as_tensor(np.ascontiguousarray(sb).copy(), **dd)
        self.inv_s2_vec = 1.0 / self.sig_vec ** 2
        self.Xb = torch.zeros(self.nctrl, dtype=torch.float64, **dd) if Xb is None else \
            torch.as_tensor(np.asarray(Xb, dtype=np.float64), **dd).reshape(-1).clone()
        m = model
        self.nu, self.nv, self.nh = (int(np.prod(s)) for s in (m.u_shape, m.v_shape, m.h_shape))
        self.ns = self.nu + self.nv + self.nh
        self.nplane = m.nx * m.ny
        x0 = [torch.as_tensor(a, dtype=torch.float64, **dd).reshape(-1) for a in x0]
        if [a.numel() for a in x0] != [self.nu, self.nv, self.nh]:
            raise VardynError("x0 must be (u, v, h) of the model's shapes")
        self.states = torch.empty((self.nseg + 1, self.ns), dtype=torch.float64, **dd)
        self.x0 = torch.cat(x0)
        # the trajectory the forward sweep keeps for the reverse sweep: every
        # step's scaled state with its two RK3 stage states ("stages", no
        # recomputation), every step's state ("full", the stage states
        # recomputed), or every ceil(sqrt(n))-th step of each segment
        # ("checkpoint", each run also recomputed once); "auto": the first
        # that takes at most 80% of the free device memory
        if tape not in ("auto", "stages", "full", "checkpoint"):
            raise ValueError("tape must be 'auto', 'stages', 'full' or 'checkpoint'")
        if tape == "auto":
            free, _ = torch.cuda.mem_get_info(self.dev)
            need = int(self.nsteps.sum()) * self.ns * 8
            tape = "stages" if 3 * need <= 0.8 * free else ("full" if need <= 0.8 * free else "checkpoint")
        self.tape_mode = tape
        # kept-state stride per segment: 0 = every step with its RK3 stage states
        self.stride = np.array([0 if tape == "stages" else 1 if tape == "full" else max(1, int(np.ceil(np.sqrt(n))))
                                for n in self.nsteps])
        kept = np.where(self.stride == 0, 3 * self.nsteps, -(-self.nsteps // np.maximum(self.stride, 1)))
        self.keep_off = np.concatenate([[0], np.cumsum(kept)]).astype(np.int64) * self.ns
        self.keep = torch.empty(int(self.keep_off[-1]), dtype=torch.float64, **dd)
        # observation table index at each segment start / the final time
        self.seg_obs = [self.H.index(self.T[c]) if self.H.is_obs_time(self.T[c]) else -1 for c in cps[:-1]]
        self.final_obs = self.H.index(self.T[cps[-1]]) if self.H.is_obs_time(self.T[cps[-1]]) else -1
        # segments starting at a basis time; the forcing each segment runs with
        self.bseg = [i for i in range(self.nseg) if cps[i] % self.dtbasis == 0]
        if not self.bseg or self.bseg[0] != 0:
            raise VardynError("the first segment must start at a basis time")
        self.bindex = {i: k for k, i in enumerate(self.bseg)}
        self.seg_basis = np.array([max(k for k, j in enumerate(self.bseg) if j <= i) for i in range(self.nseg)])
        self.bprep = basis.prepare(self.T[cps[self.bseg]] / (3600 * 24))
        nb = len(self.bseg)
        self.bfields = torch.empty((nb, m.nx, m.ny), dtype=torch.float64, **dd)
        self.Fh = torch.zeros((nb, m.nl, m.nx, m.ny), dtype=torch.float64, **dd)
        self.badj = torch.zeros((nb, m.nx, m.ny), dtype=torch.float64, **dd)
        self.lam = torch.empty(self.ns, dtype=torch.float64, **dd)
        self.lam2 = torch.empty(self.ns, dtype=torch.float64, **dd)
        self.aF = torch.empty((m.nl, m.nx, m.ny), dtype=torch.float64, **dd)
        self.adF = torch.zeros((m.nx, m.ny), dtype=torch.float64, **dd)
        self.ssh = torch.empty((m.nx, m.ny), dtype=torch.float64, **dd)
        self.adssh = torch.zeros((m.nx, m.ny), dtype=torch.float64, **dd)
        self.X = torch.empty(self.nctrl, dtype=torch.float64, **dd)
        self.G = torch.empty(self.nctrl, dtype=torch.float64, **dd)
        self.wtmp = torch.empty(self.nctrl, dtype=torch.float64, **dd)
        self.J = torch.zeros(1, dtype=torch.float64, **dd)
        self.scratch = torch.empty(_lib.VD_REDUCE_SCRATCH, dtype=torch.float64, **dd)

    # ------------------------------------------------------------ helpers
    def _s(self):
        return current_stream_ptr(self.dev)

    def _dot(self, x, y, alpha, out):
        _lib.call("vd_dot_f64", x.numel(), ptr(x), ptr(y), float(alpha), ptr(out), ptr(self.scratch), self._s())

    def _parts(self, x):
        return x[:self.nu], x[self.nu:self.nu + self.nv], x[self.nu + self.nv:]

    def _surface(self, state):
        """SSH = sum of the layer thicknesses (mod.py:3012-3016)."""
        h = self._parts(state)[2].view(self.model.nl, self.model.nx, self.model.ny)
        torch.sum(h, dim=0, out=self.ssh)
        return self.ssh

    # --------------------------------------------------------- evaluation
    def run(self, X0):
        """One cost+gradient evaluation: (J (1,), G (nctrl,)) float64 device tensors."""
        if X0.dtype != torch.float64 or X0.device != self.dev or X0.numel() != self.nctrl:
            raise VardynError(f"control must be a float64 device tensor of {self.nctrl} elements")
        X0 = X0.reshape(-1).contiguous()
        lib = _lib.load()
        s = self._s()
        m = self.model
        J, X = self.J, self.X
        J.zero_()
        # background term (tools_4Dvar.py:262-271); J accumulates (Jo + Jb) / 2
        X.copy_(self.Xb)
        if self.prec:
            X.addcmul_(self.sig_vec, X0)
            self._dot(X0, X0, 0.5, J)
        else:
            X.add_(X0)
            torch.mul(X0, self.inv_s2_vec, out=self.wtmp)
            self._dot(X0, self.wtmp, 0.5, J)
        inv_r2 = 0.5 / self.sigma_r ** 2
        # the basis fields of every basis time, then the top-layer forcing Fh = F / 86400
        self.basis.operg_prepared(self.bprep, X, self.bfields)
        torch.mul(self.bfields, 1.0 / 86400.0, out=self.Fh[:, 0])
        self.states[0].copy_(self.x0)
        for i in range(self.nseg):
            st, nxt = self.states[i], self.states[i + 1]
            if self.seg_obs[i] >= 0:
                self.H.misfit_dev(self.seg_obs[i], self._surface(st), jo=J, inv_sigma_r2=inv_r2)
            F = self.Fh[self.seg_basis[i]]
            keep = self.keep[int(self.keep_off[i]):int(self.keep_off[i + 1])]
            _lib.check(lib.vd_sw_step_keep(m._h, *(ptr(a) for a in self._parts(st)), None, None, ptr(F),
                                           *(ptr(a) for a in self._parts(nxt)), int(self.nsteps[i]),
                                           int(self.stride[i]), ptr(keep), s), "vd_sw_step_keep")
        if self.final_obs >= 0:
            self.H.misfit_dev(self.final_obs, self._surface(self.states[-1]), jo=J, inv_sigma_r2=inv_r2)

        # reverse sweep
        lam, lam2 = self.lam, self.lam2
        lam.zero_()
        self.adssh.zero_()
        self.adF.zero_()
        nl = m.nl
        if self.final_obs >= 0:
            self.H.adj_dev(self.final_obs, self.H.tables[self.final_obs].misfit, self.adssh)
        for i in reversed(range(self.nseg)):
            # the surface adjoint goes into every layer's thickness adjoint, then is cleared
            lh = self._parts(lam)[2].view(nl, m.nx, m.ny)
            lh.add_(self.adssh.unsqueeze(0))
            self.adssh.zero_()
            F = self.Fh[self.seg_basis[i]]
            keep = self.keep[int(self.keep_off[i]):int(self.keep_off[i + 1])]
            _lib.check(lib.vd_sw_step_adj_kept(m._h, None, None, ptr(F), ptr(keep), int(self.stride[i]),
                                               *(ptr(a) for a in self._parts(lam)),
                                               *(ptr(a) for a in self._parts(lam2)), None, None, ptr(self.aF),
                                               int(self.nsteps[i]), s), "vd_sw_step_adj_kept")
            lam, lam2 = lam2, lam
            # Fh = F / 86400 on the top layer: adF += aFh[0] / 86400
            self.adF.add_(self.aF[0], alpha=1.0 / 86400.0)
            if i in self.bindex:
                self.badj[self.bindex[i]].copy_(self.adF)
                self.adF.zero_()
            if self.seg_obs[i] >= 0:
                self.H.adj_dev(self.seg_obs[i], self.H.tables[self.seg_obs[i]].misfit, self.adssh)
        G = self.G
        G.zero_()
        self.basis.operg_transpose_prepared(self.bprep, self.badj, G)
        if self.prec:
            G.mul_(self.sig_vec)
            G.add_(X0)
        else:
            G.addcmul_(X0, self.inv_s2_vec)
        return J, G

    def cost_and_grad(self, X0):
        """Host interface for scipy: numpy in, (J float, G numpy) out."""
        x = torch.as_tensor(np.asarray(X0, dtype=np.float64), device=self.dev)
        J, G = self.run(x)
        return float(J[0].item()), G.cpu().numpy()
