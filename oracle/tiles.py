"""FP64 restatement of ``QgmWithTiles`` (the tiled 1.5-layer QG model).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``mapping/models/model_qg1l/jqgm.py`` of datlas-ocean/VarDyn:

* ``gaspari_cohn`` -- jqgm.py:22-53, the fifth-order piecewise rational
  correlation function of Gaspari & Cohn (1999, QJRMS 125, eq. 4.10) with
  support radius ``c`` (the argument is scaled to 2 |r| / c).
* ``overlapping_tiles`` -- jqgm.py:1098-1113: starts every
  ``tile_size - tile_overlap`` points, the last tile of each direction pushed
  back inside the domain.
* window weights and their sum over the tiles -- jqgm.py:1073-1081.
* one tile -- ``QgmWithTiles.step_tile`` (jqgm.py:1177-1214): a ``QgmOracle``
  on the tile grid with f and c the NaN-means over the tile (:1188-1189), the
  full domain's boundary / land masks sliced to the tile (:1191-1192), the
  tile's own Helmholtz operator (:1067-1071), no beta and no bathymetry term
  (the tiled ``rhs`` has none, :1185-1230), Euler or RK2 (:1262-1283).
* ``step`` -- jqgm.py:1225-1245: MDT added to h and hb on the full domain,
  every tile stepped on its slices, the tile results blended with the windows
  and divided by their sum, MDT removed.
* ``step_tgl`` / ``step_adj`` -- jqgm.py:1247-1257 (``jvp`` / ``vjp`` of
  ``step``): the blend and the slicing are linear, so the tangent blends the
  tiles' tangents, and the adjoint slices ``w * adh1 / wsum`` into each tile,
  runs the tile adjoints and adds them back where the tiles overlap.

Pinning (tests/test_oracle_tiles.py): a single tile covering the domain
reduces to ``QgmOracle`` exactly; the Gaspari-Cohn function's value 1 at 0,
zero at and beyond its support, continuity and continuous slope where its two
pieces meet, monotone decay; tiling invariants (tiles inside the domain,
covering it, stride, last tile flush with the edge); dot-product identity
between ``step_tgl`` and ``step_adj``; Taylor test of ``step_tgl`` against
``step``.
"""
from __future__ import annotations

import numpy as np

from .qg import QgmOracle, build_mask


def gaspari_cohn(r, c):
    """Gaspari-Cohn (1999) eq. 4.10 at distances ``r`` with support radius ``c``."""
    r = np.atleast_1d(np.asarray(r, dtype=np.float64))
    if c <= 0:
        return np.zeros_like(r)
    z = 2.0 * np.abs(r) / c
    out = np.zeros_like(z)
    a = z <= 1.0
    za = z[a]
    out[a] = (((-0.25 * za + 0.5) * za + 0.625) * za - 5.0 / 3.0) * za ** 2 + 1.0
    b = (z > 1.0) & (z <= 2.0)
    zb = z[b]
    out[b] = ((((zb / 12.0 - 0.5) * zb + 0.625) * zb + 5.0 / 3.0) * zb - 5.0) * zb + 4.0 - 2.0 / (3.0 * zb)
    return out


def overlapping_tiles(ny, nx, tile_size, tile_overlap):
    """(y starts, x starts) of the tiles; the tiles are their product, y outer."""
    def starts(n):
        out = []
        for s in range(0, n, tile_size - tile_overlap):
            out.append(min(max(s, 0), n - tile_size))
        return out
    return starts(ny), starts(nx)


def tile_window(tile_size, tile_overlap):
    """Separable window of one tile: Gaspari-Cohn tapers over the overlap."""
    win = np.ones(tile_size)
    g = gaspari_cohn(np.arange(tile_size), tile_overlap)[:tile_overlap]
    win[:tile_overlap] = g[::-1]
    win[tile_size - tile_overlap:] = g
    return win[:, None] * win[None, :]


class TiledQgmOracle:
    """``QgmWithTiles`` on a (ny, nx) grid: ``dx``, ``dy`` scalars, ``f`` and ``c``
    scalars or (ny, nx) arrays, ``SSH`` NaN on land (or None)."""

    def __init__(self, dx, dy, dt, shape, c, f=1e-4, g=9.81, SSH=None, Kdiffus=None,
                 time_scheme="Euler", mdt=None, tile_size=64, tile_overlap=8):
        if time_scheme not in ("Euler", "rk2"):
            raise ValueError("QgmWithTiles steps with Euler or rk2 (jqgm.py:1262-1283)")
        ny, nx = shape
        self.ny, self.nx, self.ts, self.ov = ny, nx, tile_size, tile_overlap
        self.mdt = mdt
        full = build_mask(ny, nx, SSH)
        ind0, ind12 = full == 0, (full == 1) | (full == 2)
        f = np.broadcast_to(np.asarray(f, dtype=np.float64), (ny, nx))
        c = np.broadcast_to(np.asarray(c, dtype=np.float64), (ny, nx))
        self.ys, self.xs = overlapping_tiles(ny, nx, tile_size, tile_overlap)
        self.w = tile_window(tile_size, tile_overlap)
        self.wsum = np.zeros((ny, nx))
        self.tiles = []
        for y in self.ys:
            for x in self.xs:
                sl = (slice(y, y + tile_size), slice(x, x + tile_size))
                m = QgmOracle(dx=dx, dy=dy, dt=dt, c=float(np.nanmean(c[sl])),
                              f=float(np.nanmean(f[sl])), g=g, Kdiffus=Kdiffus,
                              time_scheme=time_scheme, shape=(tile_size, tile_size))
                m.ind0 = ind0[sl].copy()
                m.ind12 = ind12[sl].copy()
                m.ind1 = (full[sl] == 1)
                m.ind2 = (full[sl] == 2)
                m.mask = np.where(m.ind0, 0, np.where(m.ind12, 2, 3)).astype(np.int8)
                self.tiles.append((sl, m))
                self.wsum[sl] += self.w

    def _blend(self, parts):
        acc = np.zeros((self.ny, self.nx))
        for (sl, _), p in zip(self.tiles, parts):
            acc[sl] += self.w * p
        return acc / self.wsum

    def step(self, h0, hb, nstep=1):
        h0 = np.asarray(h0, dtype=np.float64)
        hb = np.asarray(hb, dtype=np.float64)
        if self.mdt is not None:
            h0, hb = h0 + self.mdt, hb + self.mdt
        h1 = self._blend([m.step(h0[sl], hb[sl], nstep) for sl, m in self.tiles])
        return h1 - self.mdt if self.mdt is not None else h1

    def step_tgl(self, dh0, h0, hb, nstep=1):
        h0 = np.asarray(h0, dtype=np.float64)
        hb = np.asarray(hb, dtype=np.float64)
        if self.mdt is not None:
            h0, hb = h0 + self.mdt, hb + self.mdt
        return self._blend([m.step_tgl(dh0[sl], h0[sl], hb[sl], nstep)[0] for sl, m in self.tiles])

    def step_adj(self, adh1, h0, hb, nstep=1):
        h0 = np.asarray(h0, dtype=np.float64)
        hb = np.asarray(hb, dtype=np.float64)
        if self.mdt is not None:
            h0, hb = h0 + self.mdt, hb + self.mdt
        share = np.asarray(adh1, dtype=np.float64) / self.wsum
        adh0 = np.zeros((self.ny, self.nx))
        for sl, m in self.tiles:
            adh0[sl] += m.step_adj(self.w * share[sl], h0[sl], hb[sl], nstep)
        return adh0
