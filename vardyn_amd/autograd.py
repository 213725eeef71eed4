"""``Qgm.step`` as a differentiable torch operation.

The reference differentiates its jitted step with ``jax.jvp`` / ``jax.vjp``
(mapping/models/model_qg1l/jqgm.py:559-569).  ``QgStep`` puts the
hand-written tangent-linear and adjoint kernels of libvardyn_hip behind
``torch.autograd``:

* reverse mode (``backward``): ``vd_qg_step_adj`` -- recomputes the segment
  from ``h0`` and runs the reverse sweep, like the reference's ``vjp`` of
  ``step_jit``;
* forward mode (``jvp``, through ``torch.autograd.forward_ad`` dual tensors):
  ``vd_qg_step_tgl``.  (The ``torch.func`` transforms hand the function
  wrapper tensors without storage, which a C ABI cannot take.)

As in the reference (which differentiates with respect to ``h0`` only), the
boundary field ``hb`` is a constant: asking for its gradient raises.
"""
from __future__ import annotations

import torch

from ._lib import VardynError
from .qg import QgPlan


class QgStep(torch.autograd.Function):
    """``h1 = QgStep.apply(h0, hb, plan, nstep)``; ``hb`` may be ``None`` (zero)."""

    @staticmethod
    def forward(h0, hb, plan: QgPlan, nstep: int):
        return plan.step(h0.contiguous(), None if hb is None else hb.contiguous(), int(nstep))

    @staticmethod
    def setup_context(ctx, inputs, output):
        h0, hb, plan, nstep = inputs
        if hb is not None and hb.requires_grad:
            raise VardynError("QgStep differentiates with respect to h0 only; hb must not "
                              "require grad (jqgm.py:559-569)")
        ctx.save_for_backward(h0, hb)
        ctx.save_for_forward(h0, hb)
        ctx.plan, ctx.nstep = plan, int(nstep)

    @staticmethod
    def backward(ctx, grad_h1):
        h0, hb = ctx.saved_tensors
        adh0 = ctx.plan.step_adj(grad_h1.contiguous(), h0.contiguous(),
                                 None if hb is None else hb.contiguous(), ctx.nstep)
        return adh0, None, None, None

    @staticmethod
    def jvp(ctx, dh0, dhb, _plan, _nstep):
        if dhb is not None:
            raise VardynError("QgStep has no tangent with respect to hb")
        h0, hb = ctx.saved_tensors
        return ctx.plan.step_tgl(dh0.contiguous(), h0.contiguous(),
                                 None if hb is None else hb.contiguous(), ctx.nstep)


def qg_step(h0, hb, plan: QgPlan, nstep: int):
    """Differentiable ``Qgm.step`` (jqgm.py:517-557) on device tensors."""
    return QgStep.apply(h0, hb, plan, nstep)
