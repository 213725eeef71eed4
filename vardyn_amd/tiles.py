"""``QgmWithTiles`` on the GPU: the 1.5-layer QG model on overlapping tiles.

Reference: ``QgmWithTiles`` in mapping/models/model_qg1l/jqgm.py:1033-1258
(datlas-ocean/VarDyn).  The domain is covered by ``tile_size`` x ``tile_size``
tiles that overlap by ``tile_overlap`` points (jqgm.py:1098-1113).  Each tile
is stepped as a closed QG problem of its own -- f and c the NaN-means over the
tile (:1188-1189), the full domain's boundary and land masks sliced to the
tile (:1191-1192), the tile's own Helmholtz operator (:1067-1071), no beta and
no bathymetry term (:1185-1230), Euler or RK2 -- and the tile results are
blended with Gaspari-Cohn windows over the overlaps, divided by the summed
windows (:1073-1081, :1225-1245).  ``step_tgl`` / ``step_adj`` are the
reference's ``jvp`` / ``vjp`` of ``step`` (:1247-1257): tangent-linear and
adjoint of every tile composed with the (linear) slicing and blending.

Here every tile is a plan of libvardyn_hip (its own constants and mask), and
the tile plans form a ragged group (``vd_group_*``): every RK stage, Helmholtz
inversion and adjoint stage of a step is one launch over all tiles, so a step
costs the launches of one problem however many tiles there are (the reference
default, 32-point tiles overlapping by 16, gives 1024 tiles at 512 x 512).
Segment start and end are group launches too (``vd_group_prep/finish/
adj_init/adj_final``); the tangent-linear model runs tile by tile through the
plans, and ``engine="plans"`` steps every tile through its own plan.  Slicing, blending and their transposes are the
``vd_tiles_*`` kernels.  Mask codes of a tile: 0 land, 2 where the
domain's mask is 1 or 2 (q from hb, no tendency), and for the domain's
interior points 3 (active) in tile rows and columns 2..n-3, where the tiled
``rhs`` evaluates the advection (jqgm.py:377-396 on the tile), 4 (q from the
Laplacian, no tendency) in rows and columns 1 and n-2, and 5 (q = 0, no
tendency) on the tile's ring, whose PV the tiled h2pv leaves at zero and the
advection of rows 2 and n-3 reads.
"""
from __future__ import annotations


import numpy as np
import torch

from . import _lib
from ._lib import VardynError
from .qg import QgPlan, build_mask, current_stream_ptr, ptr
from .ragged import QgGroup


def gaspari_cohn(r, c):
    """Gaspari & Cohn (1999, QJRMS 125, eq. 4.10) at distances ``r`` with
    support radius ``c`` (jqgm.py:22-53)."""
    r = np.atleast_1d(np.asarray(r, dtype=np.float64))
    if c <= 0:
        return np.zeros_like(r)
    z = 2.0 * np.abs(r) / c
    out = np.zeros_like(z)
    near = z <= 1.0
    zn = z[near]
    out[near] = -0.25 * zn ** 5 + 0.5 * zn ** 4 + 0.625 * zn ** 3 - 5.0 / 3.0 * zn ** 2 + 1.0
    far = (z > 1.0) & (z <= 2.0)
    zf = z[far]
    out[far] = (zf ** 5 / 12.0 - 0.5 * zf ** 4 + 0.625 * zf ** 3 + 5.0 / 3.0 * zf ** 2 - 5.0 * zf
                + 4.0 - 2.0 / (3.0 * zf))
    return out


def tile_starts(n, tile_size, tile_overlap):
    """Tile starts along one axis (jqgm.py:1098-1113)."""
    if not 0 <= tile_overlap < tile_size <= n:
        raise VardynError("need 0 <= tile_overlap < tile_size <= domain size")
    return [min(s, n - tile_size) for s in range(0, n, tile_size - tile_overlap)]


def tile_mask(full_mask, y, x, ts):
    """Mask codes of the tile at (y, x) from the domain's mask codes."""
    m = np.asarray(full_mask)[y:y + ts, x:x + ts]
    out = np.full((ts, ts), 3, dtype=np.int8)
    inner = np.zeros((ts, ts), dtype=bool)
    inner[1:-1, 1:-1] = True
    band = np.zeros((ts, ts), dtype=bool)
    band[2:-2, 2:-2] = True
    out[inner & ~band] = 4
    out[~inner] = 5
    out[(m == 1) | (m == 2)] = 2
    out[m == 0] = 0
    return out


class QgmWithTiles:
    """Drop-in for ``QgmWithTiles`` (jqgm.py:1033-1258): the reference
    constructor arguments and the ``step`` / ``step_tgl`` / ``step_adj``
    calls (and their ``*_jit`` names) on numpy or torch arrays."""

    def __init__(self, dx=None, dy=None, dt=None, SSH=None, c=None, Kdiffus=None, upwind=3,
                 g=9.81, f=1e-4, time_scheme="Euler", mdt=None, tile_size=64, tile_overlap=8,
                 precision="float64", device=None, dst_split=0, engine="group", **kwargs):
        if upwind != 3:
            raise VardynError("only the 3rd-order upwind scheme exists in the reference")
        if time_scheme not in ("Euler", "rk2"):
            raise VardynError("QgmWithTiles steps with Euler or rk2 (jqgm.py:1262-1283)")
        if engine not in ("group", "plans"):
            raise VardynError("engine must be 'group' (shared launches) or 'plans' (tile by tile)")
        self.engine = engine
        if hasattr(dx, "__len__"):
            ny, nx = np.shape(dx)
            dxs = float(np.asarray(dx)[ny // 2, nx // 2])
            dys = float(np.asarray(dy)[ny // 2, nx // 2])
        else:
            if SSH is None:
                raise VardynError("give 2-D dx/dy or an SSH field for the grid shape")
            ny, nx = np.shape(SSH)
            dxs, dys = float(dx), float(dy)
        self.ny, self.nx, self.ts, self.ov = ny, nx, int(tile_size), int(tile_overlap)
        self.dt = float(dt)
        self.time_scheme = time_scheme
        ssh = None if SSH is None else np.asarray(SSH, dtype=np.float64)
        full = build_mask(ny, nx, ssh)
        f = np.broadcast_to(np.asarray(f, dtype=np.float64), (ny, nx))
        c = np.broadcast_to(np.asarray(c, dtype=np.float64), (ny, nx))
        self.ys = tile_starts(ny, self.ts, self.ov)
        self.xs = tile_starts(nx, self.ts, self.ov)
        win = np.ones(self.ts)
        gc = gaspari_cohn(np.arange(self.ts), self.ov)[:self.ov]
        win[:self.ov] = gc[::-1]
        win[self.ts - self.ov:] = gc
        w = win[:, None] * win[None, :]
        wsum = np.zeros((ny, nx))
        self.plans = []
        for y in self.ys:
            for x in self.xs:
                sl = (slice(y, y + self.ts), slice(x, x + self.ts))
                wsum[sl] += w
                self.plans.append(QgPlan(self.ts, self.ts, dxs, dys, self.dt,
                                         float(np.nanmean(c[sl])), float(np.nanmean(f[sl])), g=g,
                                         time_scheme=time_scheme, precision=precision,
                                         kdiffus=Kdiffus, mask=tile_mask(full, y, x, self.ts),
                                         device=device, dst_split=dst_split))
        p0 = self.plans[0]
        self.dtype, self.device, self.prec = p0.dtype, p0.device, p0.prec
        self.group = None
        if engine == "group":
            self.group = QgGroup(self.plans)
            tsq = self.ts * self.ts
            if self.group.offsets != [t * tsq for t in range(self.ntile + 1)]:
                raise VardynError("tile plans must lie back to back in the group's flat fields")
        dd = dict(device=self.device)
        self._ys = torch.tensor(self.ys, dtype=torch.int32, **dd)
        self._xs = torch.tensor(self.xs, dtype=torch.int32, **dd)
        self._w = torch.as_tensor(w, dtype=self.dtype, **dd).contiguous()
        self._wsum = torch.as_tensor(wsum, dtype=self.dtype, **dd).contiguous()
        self._mdt = None if mdt is None else torch.as_tensor(np.asarray(mdt, dtype=np.float64),
                                                            dtype=self.dtype, **dd).contiguous()
        self.step_jit = self.step
        self.step_tgl_jit = self.step_tgl
        self.step_adj_jit = self.step_adj

    def autotune(self, reps=5):
        """Pick the tile group's adjoint tile shape and inversion form (fused per
        tile or GEMM stages) by timing them (``vd_group_autotune``); returns
        the choices, or None for ``engine="plans"``."""
        return None if self.group is None else self.group.autotune(reps)

    # ------------------------------------------------------------- helpers
    @property
    def ntile(self):
        return len(self.plans)

    def _dims(self):
        return (self.prec, self.ny, self.nx, self.ts, len(self.ys), len(self.xs),
                ptr(self._ys), ptr(self._xs))

    def _field(self, a):
        t = torch.as_tensor(a).to(device=self.device, dtype=self.dtype).contiguous()
        if tuple(t.shape) != (self.ny, self.nx):
            raise VardynError(f"expected a ({self.ny}, {self.nx}) field, got {tuple(t.shape)}")
        return t

    def _stack(self):
        return torch.empty((self.ntile, self.ts, self.ts), dtype=self.dtype, device=self.device)

    def _gather(self, full, add=None):
        out = self._stack()
        _lib.call("vd_tiles_gather", *self._dims(), ptr(full), ptr(add), ptr(out),
                  current_stream_ptr(self.device))
        return out

    def _blend(self, flat, sub=None):
        out = torch.empty((self.ny, self.nx), dtype=self.dtype, device=self.device)
        _lib.call("vd_tiles_blend", *self._dims(), ptr(self._w), ptr(self._wsum), ptr(flat),
                  ptr(sub), ptr(out), current_stream_ptr(self.device))
        return out

    @staticmethod
    def _like_input(x, out):
        return out.cpu().numpy() if isinstance(x, np.ndarray) else out

    def _s(self):
        return current_stream_ptr(self.device)

    def _prep(self, H0, HB):
        """Segment start of every tile: (h, q, q_b) = (h0, h2pv(h0, hb), h2pv(hb, hb))."""
        hc, qc, qb = self._stack(), self._stack(), self._stack()
        self.group.prep(self.ntile, ptr(H0), ptr(HB), ptr(hc), ptr(qc), None, ptr(qb))
        return hc, qc, qb

    def _forward_group(self, H0, HB, nstep, tape=None):
        """``nstep`` steps of every tile in shared launches (jqgm.py:398-503 per
        tile, as ``vd_qg_step`` does it); returns the last (h, q) and, with a
        ``tape`` of shape (nstep + 1, stages, 2, ntile, ts, ts), keeps h and q
        of every stage there."""
        g, n, dt = self.group, self.ntile, self.dt
        hc, qc, qb = self._prep(H0, HB)
        hb, qbp = ptr(HB), ptr(qb)
        if tape is not None:
            tape[0, 0, 0].copy_(hc)
            tape[0, 0, 1].copy_(qc)
            hc, qc = tape[0, 0, 0], tape[0, 0, 1]
        bufs = [(hc, qc), (self._stack(), self._stack())]
        hs1, qs1 = self._stack(), self._stack()
        for k in range(int(nstep)):
            if tape is not None:
                hn, qn = tape[k + 1, 0, 0], tape[k + 1, 0, 1]
                if self.time_scheme == "rk2":
                    hs1, qs1 = tape[k, 1, 0], tape[k, 1, 1]
            else:
                hn, qn = bufs[(k + 1) % 2]
            if self.time_scheme == "Euler":
                g.stage(n, ptr(hc), ptr(qc), ptr(qc), None, None, 0, 0, dt, ptr(qn))
            else:
                g.stage(n, ptr(hc), ptr(qc), ptr(qc), None, None, 0, 0, 0.5 * dt, ptr(qs1))
                g.inversion(n, ptr(qs1), ptr(hs1), Fsub=qbp, hb=hb)
                g.stage(n, ptr(hs1), ptr(qs1), ptr(qc), None, None, 0, 0, dt, ptr(qn))
            g.inversion(n, ptr(qn), ptr(hn), Fsub=qbp, hb=hb)
            hc, qc = hn, qn
        return hc, qc

    # ------------------------------------------------------------- model calls
    def step(self, h0, hb, nstep=1):
        """``QgmWithTiles.step`` (jqgm.py:1225-1245)."""
        h, b = self._field(h0), self._field(hb)
        H0, HB = self._gather(h, self._mdt), self._gather(b, self._mdt)
        H1 = self._stack()
        if self.group is not None:
            hc, _ = self._forward_group(H0, HB, nstep)
            self.group.finish(self.ntile, ptr(hc), ptr(H1))
        else:
            for t, plan in enumerate(self.plans):
                plan.step(H0[t], HB[t], int(nstep), h1=H1[t])
        return self._like_input(h0, self._blend(H1, self._mdt))

    def step_tgl(self, dh0, h0, hb, nstep=1):
        """``jvp`` of ``step`` (jqgm.py:1247-1251)."""
        d, h, b = self._field(dh0), self._field(h0), self._field(hb)
        D0, H0, HB = self._gather(d), self._gather(h, self._mdt), self._gather(b, self._mdt)
        D1 = self._stack()
        for t, plan in enumerate(self.plans):
            plan.step_tgl(D0[t], H0[t], HB[t], int(nstep), dh1=D1[t])
        return self._like_input(dh0, self._blend(D1))

    def step_adj(self, adh1, h0, hb, nstep=1):
        """``vjp`` of ``step`` (jqgm.py:1253-1257)."""
        a, h, b = self._field(adh1), self._field(h0), self._field(hb)
        A1 = self._stack()
        _lib.call("vd_tiles_blend_adj", *self._dims(), ptr(self._w), ptr(self._wsum), ptr(a),
                  ptr(A1), current_stream_ptr(self.device))
        H0, HB = self._gather(h, self._mdt), self._gather(b, self._mdt)
        A0 = self._stack()
        if self.group is not None:
            self._adjoint_group(A1, H0, HB, int(nstep), A0)
        else:
            tape = self.plans[0].new_tape(int(nstep))  # the tiles share one tape, one after another
            for t, plan in enumerate(self.plans):
                plan.step_adj(A1[t], H0[t], HB[t], int(nstep), adh0=A0[t], tape=tape)
        out = torch.empty((self.ny, self.nx), dtype=self.dtype, device=self.device)
        _lib.call("vd_tiles_gather_adj", *self._dims(), ptr(A0), ptr(out),
                  current_stream_ptr(self.device))
        return self._like_input(adh1, out)

    def _adjoint_group(self, A1, H0, HB, nstep, A0):
        """Adjoint of ``nstep`` steps of every tile in shared launches (the
        reverse sweep of ``vd_qg_step_adj``): forward recompute onto a tape,
        then per step the adjoint stages and inversions."""
        g, n, dt = self.group, self.ntile, self.dt
        nst = 1 if self.time_scheme == "Euler" else 2
        tape = torch.empty((nstep + 1, nst, 2, self.ntile, self.ts, self.ts), dtype=self.dtype,
                           device=self.device)
        self._forward_group(H0, HB, nstep, tape=tape)
        lh, lq, lqx, lq1, lhs = (self._stack() for _ in range(5))
        g.adj_init(n, ptr(A1), ptr(lh), ptr(lq))
        for k in reversed(range(nstep)):
            # h_{k+1} = pv2h(q_{k+1}): lq += P lh
            g.inversion(n, ptr(lh), ptr(lq), accumulate=True)
            h0k, q0k = ptr(tape[k, 0, 0]), ptr(tape[k, 0, 1])
            if self.time_scheme == "Euler":
                g.stage_adj(n, h0k, q0k, ptr(lq), None, None, dt, 0, 0, ptr(lq), None, None,
                            ptr(lqx), ptr(lh))
                lq, lqx = lqx, lq
            else:
                g.stage_adj(n, ptr(tape[k, 1, 0]), ptr(tape[k, 1, 1]), ptr(lq), None, None, dt, 0,
                            0, None, None, None, ptr(lq1), ptr(lhs))
                g.inversion(n, ptr(lhs), ptr(lq1), accumulate=True)
                g.stage_adj(n, h0k, q0k, ptr(lq1), None, None, 0.5 * dt, 0, 0, ptr(lq), ptr(lq1),
                            None, ptr(lq), ptr(lh))
        g.adj_final(n, ptr(tape[0, 0, 0]), ptr(lh), ptr(lq), ptr(A0))
