"""Drop-in replacement for ``mapping/models/model_qg1l/jqgm.py`` (class ``Qgm``).

``Model_qg1l_jax`` loads ``f'{config.MOD.dir_model}/jqgm.py'`` with
``SourceFileLoader`` and instantiates ``getattr(module, config.MOD.name_class)``
(mapping/src/mod.py:479-486).  Pointing ``dir_model`` at this package's
directory makes it pick up this ``Qgm``: same constructor (jqgm.py:113),
same reductions of its arguments, and the same ``step_jit`` /
``step_tgl_jit`` / ``step_adj_jit`` callables (jqgm.py:223-236), computed by
the HIP kernels of libvardyn_hip.

Arrays may be numpy (results are returned as numpy) or torch tensors
(results stay on the device).  Precision follows ``USE_FLOAT64`` of the
reference (mapping/src/config.py:6-8): float32 unless ``precision='float64'``
is given or the environment variable ``VARDYN_FLOAT64=1`` is set.

``QgmWithTiles`` (jqgm.py:1033-1258, ``name_class='QgmWithTiles'``) is the
tiled model of ``vardyn_amd.tiles``; it computes in FP64 unless
``precision='float32'`` is given.

This module uses absolute imports only, so it also works when loaded by
path as a top-level module.
"""
from __future__ import annotations

import os

import numpy as np
import torch

from vardyn_amd.qg import QgPlan, build_mask
from vardyn_amd.tiles import QgmWithTiles  # noqa: F401  (name_class='QgmWithTiles')


def _default_precision():
    return "float64" if os.environ.get("VARDYN_FLOAT64", "0") == "1" else "float32"


class Qgm:
    """1.5-layer QG model (jqgm.py:108-569) on the GPU."""

    def __init__(self, dx=None, dy=None, dt=None, SSH=None, c=None, Kdiffus=None, upwind=3,
                 g=9.81, f=1e-4, time_scheme="Euler", compile=True, mdt=None,
                 bathymetry_PV_term=None, precision=None, device=None, **kwargs):
        if upwind != 3:
            raise NotImplementedError("the reference advects PV with the 3rd-order upwind "
                                      "scheme only (jqgm.py:377-396)")
        precision = precision or _default_precision()
        npdt = np.float64 if precision == "float64" else np.float32
        self.dtype = "float64" if precision == "float64" else "float32"
        ny, nx = np.shape(dx)
        self.ny, self.nx = ny, nx
        # Grid spacing from the centre pixel (jqgm.py:125-131)
        if hasattr(dx, "__len__"):
            self.dx = npdt(np.asarray(dx)[int(ny / 2), int(nx / 2)])
            self.dy = npdt(np.asarray(dy)[int(ny / 2), int(nx / 2)])
        else:
            self.dx, self.dy = dx, dy
        self.dt = dt
        self.g = g
        # Coriolis: nanmean, and a beta plane when f is an array (jqgm.py:139-146)
        if hasattr(f, "__len__"):
            f = np.asarray(f, dtype=np.float64)
            self.f = npdt(np.nanmean(f))
            self.beta = (f[2:, :] - f[:-2, :]) / (2 * float(self.dy))
        else:
            self.f = npdt(f)
            self.beta = None
        self.c = npdt(np.nanmean(c)) if hasattr(c, "__len__") else npdt(c)
        self.mdt = mdt
        self.bathymetry_PV_term = bathymetry_PV_term
        self.upwind = upwind
        self.time_scheme = time_scheme
        self.Kdiffus = Kdiffus
        self.mask = build_mask(ny, nx, SSH)
        self.ind0 = self.mask == 0
        self.ind1 = self.mask == 1
        self.ind2 = self.mask == 2
        self.ind12 = self.ind1 + self.ind2

        fbathy = None
        if bathymetry_PV_term is not None:
            fbathy = float(self.f) * np.asarray(bathymetry_PV_term, dtype=np.float64)
        self.plan = QgPlan(ny, nx, float(self.dx), float(self.dy), float(dt), float(self.c),
                           float(self.f), g=g, time_scheme=time_scheme, precision=self.dtype,
                           kdiffus=Kdiffus, mask=self.mask,
                           beta=self.beta, fbathy=fbathy,
                           mdt=None if mdt is None else np.asarray(mdt, dtype=np.float64),
                           device=device)
        self._tape = None
        self._tape_steps = -1
        # the reference's jitted callables (jqgm.py:223-236)
        self.step_jit = self.step
        self.step_tgl_jit = self.step_tgl
        self.step_adj_jit = self.step_adj
        self.h2pv_jit = self.h2pv
        self.pv2h_jit = self.pv2h

    # ------------------------------------------------------------------
    def _in(self, a):
        return self.plan.field(a)

    @staticmethod
    def _out(t, like):
        if isinstance(like, torch.Tensor):
            return t
        return t.cpu().numpy()

    def _tape_for(self, nstep):
        if self._tape is None or self._tape_steps < nstep:
            self._tape = self.plan.new_tape(nstep)
            self._tape_steps = nstep
        return self._tape

    def step(self, h0, hb, nstep=1):
        """Qgm.step (jqgm.py:517-557)."""
        h1 = self.plan.step(self._in(h0), self._in(hb), int(nstep))
        return self._out(h1, h0)

    def step_tgl(self, dh0, h0, hb, nstep=1):
        """Qgm.step_tgl (jqgm.py:559-563)."""
        dh1 = self.plan.step_tgl(self._in(dh0), self._in(h0), self._in(hb), int(nstep))
        return self._out(dh1, dh0)

    def step_adj(self, adh0, h0, hb, nstep=1):
        """Qgm.step_adj (jqgm.py:565-569)."""
        tape = self._tape_for(int(nstep))
        ad = self.plan.step_adj(self._in(adh0), self._in(h0), self._in(hb), int(nstep), tape=tape)
        return self._out(ad, adh0)

    def h2pv(self, h, hb):
        """Qgm.h2pv (jqgm.py:264-293)."""
        return self._out(self.plan.h2pv(self._in(h), self._in(hb)), h)

    def pv2h(self, q, hb, qb):
        """Qgm.pv2h (jqgm.py:295-319)."""
        return self._out(self.plan.pv2h(self._in(q), self._in(hb), self._in(qb)), q)
