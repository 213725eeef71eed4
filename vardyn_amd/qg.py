"""Device-side QG model plan: owns a ``vd_plan`` and calls the C ABI on torch
device tensors (torch provides device memory and the current HIP stream).

Mirrors the numerics of ``Qgm`` (mapping/models/model_qg1l/jqgm.py:108-569,
datlas-ocean/VarDyn); see include/vardyn.h for the contract of each call.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch

from . import _lib
from ._lib import QgParams, VardynError


def build_mask(ny, nx, SSH=None):
    """Mask of jqgm.py:179-217: 3 active, 1 border/coast, 2 next to them, 0 land."""
    mask = 3 * np.ones((ny, nx), dtype=np.int8)
    mask[0, :] = 1
    mask[:, 0] = 1
    mask[-1, :] = 1
    mask[:, -1] = 1
    mask[1, 1:-1] = 2
    mask[1:-1, 1] = 2
    mask[-2, 1:-1] = 2
    mask[-3, 1:-1] = 2
    mask[1:-1, -2] = 2
    mask[1:-1, -3] = 2
    if SSH is not None:
        land = np.isnan(np.asarray(SSH, dtype=np.float64))
        mask[land] = 0
        for i, j in np.argwhere(land):
            for p1 in range(-2, 3):
                for p2 in range(-2, 3):
                    it, jt = i + p1, j + p2
                    if 0 <= it < ny and 0 <= jt < nx:
                        if mask[it, jt] >= 2 and abs(p1) <= 1 and abs(p2) <= 1:
                            mask[it, jt] = 1
                        elif mask[it, jt] == 3:
                            mask[it, jt] = 2
    return mask


def _torch_dtype(precision):
    if precision in ("float64", "f64", torch.float64, np.float64, _lib.VD_F64):
        return torch.float64
    if precision in ("float32", "f32", torch.float32, np.float32, _lib.VD_F32):
        return torch.float32
    raise ValueError(f"unsupported precision {precision!r}")


def current_stream_ptr(device=None):
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class QgPlan:
    """One QG problem on one GPU (constants + scratch in device memory)."""

    def __init__(self, ny, nx, dx, dy, dt, c, f0, g=9.81, time_scheme="rk2",
                 precision="float64", kdiffus=None, mask=None, beta=None, fbathy=None,
                 mdt=None, device=None, dst_dense=None, gemm_config=None, nbatch=1,
                 dst_split=None):
        if not torch.cuda.is_available():
            raise VardynError("no HIP device available: the QG plan runs on the GPU only")
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None
                                   else torch.device(device).index or 0)
        self.ny, self.nx = int(ny), int(nx)
        self.nb = int(nbatch)
        if self.nb < 1:
            raise VardynError("nbatch must be >= 1")
        self.dtype = _torch_dtype(precision)
        self.prec = _lib.VD_F64 if self.dtype == torch.float64 else _lib.VD_F32
        if time_scheme not in _lib.SCHEMES:
            raise ValueError(f"time_scheme must be one of {list(_lib.SCHEMES)}")
        self.time_scheme = time_scheme
        self.nstage = {"Euler": 1, "rk2": 2, "rk3": 3}[time_scheme]
        self.dt = float(dt)
        self._keep = []

        def host(a, shape, dtype):
            if a is None:
                return None
            a = np.ascontiguousarray(np.asarray(a, dtype=dtype))
            if a.shape != shape:
                raise VardynError(f"expected shape {shape}, got {a.shape}")
            self._keep.append(a)
            return C.c_void_p(a.ctypes.data)

        p = QgParams()
        p.ny, p.nx = self.ny, self.nx
        p.precision = self.prec
        p.time_scheme = _lib.SCHEMES[time_scheme]
        p.dx, p.dy, p.dt = float(dx), float(dy), float(dt)
        c_arr = np.atleast_1d(np.asarray(c, dtype=np.float64))
        if c_arr.size not in (1, self.nb):
            raise VardynError("c must be a scalar or hold one value per batch member")
        p.g, p.f0, p.c = float(g), float(f0), float(c_arr[0])
        self.c = c_arr if c_arr.size == self.nb else np.full(self.nb, c_arr[0])
        p.nbatch = self.nb
        p.c_batch = host(self.c, (self.nb,), np.float64)
        p.has_kdiffus = 0 if kdiffus is None else 1
        p.kdiffus = 0.0 if kdiffus is None else float(kdiffus)
        p.mask = host(mask, (self.ny, self.nx), np.int8)
        p.beta = host(beta, (self.ny - 2, self.nx), np.float64)
        p.fbathy = host(fbathy, (self.ny, self.nx), np.float64)
        p.mdt = host(mdt, (self.ny, self.nx), np.float64)
        if dst_dense is None:
            # VARDYN_DST_DENSE=1 selects the full sine matrices for A/B comparisons
            dst_dense = os.environ.get("VARDYN_DST_DENSE", "0") == "1"
        p.dst_dense = 1 if dst_dense else 0
        self.folded = (not dst_dense) and (self.ny - 2) % 2 == 0 and (self.nx - 2) % 2 == 0
        if gemm_config is None:
            # VARDYN_GEMM_CONFIG=k forces tile configuration k-1 (A/B comparisons)
            gemm_config = int(os.environ.get("VARDYN_GEMM_CONFIG", "0"))
        p.gemm_config = int(gemm_config)
        if dst_split is None:
            # VARDYN_DST_SPLIT=n selects the split-product GEMMs with n digits
            dst_split = int(os.environ.get("VARDYN_DST_SPLIT", "0"))
        p.dst_split = int(dst_split)
        self.dst_split = p.dst_split
        self.params = p
        handle = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.call("vd_qg_plan_create", C.byref(p), C.byref(handle))
        self._h = handle
        self._keep.clear()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().vd_qg_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    # ---------------------------------------------------------------- helpers
    @property
    def shape(self):
        """Shape of one field argument: (ny, nx), or (nbatch, ny, nx) for batches."""
        return (self.ny, self.nx) if self.nb == 1 else (self.nb, self.ny, self.nx)

    @property
    def _numel(self):
        return self.nb * self.ny * self.nx

    def tape_bytes(self, nstep):
        b = C.c_size_t()
        _lib.call("vd_qg_tape_bytes", self._h, int(nstep), C.byref(b))
        return b.value

    def autotune(self, reps=10):
        """Time the GEMM tile configurations on the device and keep the fastest;
        returns the chosen configuration index of each inversion GEMM."""
        chosen = (C.c_int * 4)()
        _lib.call("vd_qg_autotune", self._h, int(reps), chosen, current_stream_ptr(self.device))
        return list(chosen)

    def plan_bytes(self):
        b = C.c_size_t()
        _lib.call("vd_qg_plan_bytes", self._h, C.byref(b))
        return b.value

    def new_tape(self, nstep):
        n = self.tape_bytes(nstep) // (8 if self.prec == _lib.VD_F64 else 4)
        return torch.empty(n, dtype=self.dtype, device=self.device)

    def empty(self):
        return torch.empty(self.shape, dtype=self.dtype, device=self.device)

    def field(self, a):
        """Device tensor of the plan dtype/shape from numpy or torch input."""
        if isinstance(a, torch.Tensor):
            t = a.to(device=self.device, dtype=self.dtype)
        else:
            t = torch.as_tensor(np.asarray(a), device=self.device).to(self.dtype)
        if tuple(t.shape) != self.shape:
            raise VardynError(f"field shape {tuple(t.shape)} != plan shape {self.shape}")
        return t.contiguous()

    def _check(self, *ts):
        for t in ts:
            if t is None:
                continue
            if (t.device != self.device or t.dtype != self.dtype or not t.is_contiguous()):
                raise VardynError("fields must be contiguous device tensors of the plan dtype")
            if t.numel() != self._numel:
                raise VardynError(f"field has {t.numel()} elements, the plan expects "
                                  f"{self._numel} {tuple(self.shape)}")

    def _check_tape(self, tape, nstep):
        need = self.tape_bytes(nstep)
        if tape.dtype != self.dtype or tape.device != self.device or not tape.is_contiguous():
            raise VardynError("tape must be a contiguous device tensor of the plan dtype")
        if tape.numel() * tape.element_size() < need:
            raise VardynError(f"tape holds {tape.numel() * tape.element_size()} bytes, "
                              f"{need} needed for {nstep} steps")

    # ------------------------------------------------------------------- calls
    def step(self, h0, hb, nstep, h1=None, tape=None):
        h1 = self.empty() if h1 is None else h1
        self._check(h0, hb, h1)
        if tape is not None:
            self._check_tape(tape, nstep)
        _lib.call("vd_qg_step", self._h, ptr(h0), ptr(hb), ptr(h1), int(nstep), ptr(tape),
                  current_stream_ptr(self.device))
        return h1

    def step_tgl(self, dh0, h0, hb, nstep, dh1=None, h1=None):
        dh1 = self.empty() if dh1 is None else dh1
        self._check(dh0, h0, hb, dh1, h1)
        _lib.call("vd_qg_step_tgl", self._h, ptr(dh0), ptr(h0), ptr(hb), ptr(dh1), ptr(h1),
                  int(nstep), current_stream_ptr(self.device))
        return dh1

    def step_adj(self, adh1, h0, hb, nstep, adh0=None, tape=None, tape_valid=False):
        adh0 = self.empty() if adh0 is None else adh0
        if tape is None:
            if tape_valid:
                raise VardynError("tape_valid requires a tape")
            tape = self.new_tape(nstep)
        self._check(adh1, h0, hb, adh0)
        self._check_tape(tape, nstep)
        _lib.call("vd_qg_step_adj", self._h, ptr(adh1), ptr(h0), ptr(hb), ptr(adh0), int(nstep),
                  ptr(tape), 1 if tape_valid else 0, current_stream_ptr(self.device))
        return adh0

    def pv2h(self, q, hb=None, qb=None, out=None):
        out = self.empty() if out is None else out
        self._check(q, hb, qb, out)
        _lib.call("vd_qg_pv2h", self._h, ptr(q), ptr(hb), ptr(qb), ptr(out),
                  current_stream_ptr(self.device))
        return out

    def h2pv(self, h, hb, out=None):
        out = self.empty() if out is None else out
        self._check(h, hb, out)
        _lib.call("vd_qg_h2pv", self._h, ptr(h), ptr(hb), ptr(out),
                  current_stream_ptr(self.device))
        return out

    def rhs(self, h, q, base=None, coef=1.0, out=None):
        """``base + coef * rhs(h, q)`` (Qgm.rhs, jqgm.py:321-375, fused with one
        explicit stage update; ``base=None`` means zero)."""
        out = self.empty() if out is None else out
        self._check(h, q, base, out)
        _lib.call("vd_qg_rhs", self._h, ptr(h), ptr(q), ptr(base), float(coef), ptr(out),
                  current_stream_ptr(self.device))
        return out

    def rhs_adj(self, h, q, lam, coef=1.0, acc=None, lq=None, lh=None):
        """Transpose of ``rhs`` at (h, q) applied to ``lam`` (the vjp of
        ``coef * rhs(h2uv(h), q)``, jqgm.py:238-262, 321-375): returns
        ``(acc + coef * drhs/dq^T lam, coef * drhs/dh^T lam)``."""
        lq = self.empty() if lq is None else lq
        lh = self.empty() if lh is None else lh
        self._check(h, q, lam, acc, lq, lh)
        _lib.call("vd_qg_rhs_adj", self._h, ptr(h), ptr(q), ptr(lam), float(coef), ptr(acc),
                  ptr(lq), ptr(lh), current_stream_ptr(self.device))
        return lq, lh
