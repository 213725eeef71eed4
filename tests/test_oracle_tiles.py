"""The tiled QG oracle (QgmWithTiles, jqgm.py:1033-1258), pinned without the
reference: the Gaspari-Cohn function by its defining properties, the tiling
by its invariants, one tile covering the domain by the untiled oracle, and the
tangent-linear / adjoint pair by the dot-product identity and a Taylor test."""
import numpy as np
import pytest

from oracle.qg import QgmOracle
from oracle.tiles import TiledQgmOracle, gaspari_cohn, overlapping_tiles, tile_window


def test_gaspari_cohn_properties():
    c = 8.0
    assert gaspari_cohn(0.0, c)[0] == 1.0
    assert abs(gaspari_cohn(c, c)[0]) < 1e-15  # the far piece vanishes at the support radius
    assert np.all(gaspari_cohn(np.array([1.2 * c, 3 * c]), c) == 0.0)
    # the two pieces meet at r = c/2 with equal value and slope, and the far piece reaches 0 at c
    lo, hi = gaspari_cohn(np.array([c / 2 - 1e-10, c / 2 + 1e-10]), c)
    assert abs(lo - hi) < 1e-9
    e = 1e-7
    s1 = (gaspari_cohn(c / 2 - e, c)[0] - gaspari_cohn(c / 2 - 3 * e, c)[0]) / (2 * e)
    s2 = (gaspari_cohn(c / 2 + 3 * e, c)[0] - gaspari_cohn(c / 2 + e, c)[0]) / (2 * e)
    assert abs(s1 - s2) < 1e-5
    assert gaspari_cohn(c - e, c)[0] < 1e-12
    r = np.linspace(0, c, 401)
    g = gaspari_cohn(r, c)
    assert np.all(np.diff(g) <= 1e-15) and np.all(g >= -1e-15)
    assert np.allclose(gaspari_cohn(-r, c), g)


@pytest.mark.parametrize("n,ts,ov", [(100, 40, 8), (64, 64, 8), (130, 32, 4), (97, 20, 6)])
def test_tiling_invariants(n, ts, ov):
    ys, xs = overlapping_tiles(n, n + 3, ts, ov)
    for starts, m in ((ys, n), (xs, n + 3)):
        assert starts[0] == 0 and starts[-1] == m - ts
        assert all(0 <= s <= m - ts for s in starts)
        d = np.diff(starts)
        # stride ts - ov until the clamped tail; loop values past n - ts all clamp to the last
        # start, so the reference repeats that tile (jqgm.py:1098-1113)
        k = int(np.argmax(d != ts - ov)) if np.any(d != ts - ov) else len(d)
        assert np.all(d[:k] == ts - ov) and np.all((d[k:] >= 0) & (d[k:] < ts - ov))
        assert np.all(d[k + 1:] == 0)
        cover = np.zeros(m, dtype=int)
        for s in starts:
            cover[s:s + ts] += 1
        assert np.all(cover >= 1)
    w = tile_window(ts, ov)
    assert np.all(w > 0) and w.max() == 1.0 and np.allclose(w, w.T)


def _fields(ny, nx, seed, land=False):
    rng = np.random.default_rng(seed)
    h0 = 0.05 * rng.standard_normal((ny, nx))
    hb = 0.01 * rng.standard_normal((ny, nx))
    ssh = None
    if land:
        ssh = np.zeros((ny, nx))
        ssh[ny // 2, nx // 3] = np.nan
        ssh[ny // 3, nx // 2:nx // 2 + 2] = np.nan
        h0[np.isnan(ssh)] = np.nan
    return h0, hb, ssh, rng


@pytest.mark.parametrize("scheme", ["Euler", "rk2"])
def test_single_tile_is_the_untiled_model(scheme):
    n = 36
    h0, hb, ssh, _ = _fields(n, n, 1, land=True)
    tiled = TiledQgmOracle(10e3, 10e3, 1200.0, (n, n), c=2.7, f=1e-4, SSH=ssh, Kdiffus=30.0,
                           time_scheme=scheme, tile_size=n, tile_overlap=6)
    # tile_size == n: every start clamps to 0, so the reference repeats the one tile
    assert all(sl == (slice(0, n), slice(0, n)) for sl, _ in tiled.tiles)
    plain = QgmOracle(dx=10e3, dy=10e3, dt=1200.0, c=2.7, f=1e-4, SSH=ssh, Kdiffus=30.0,
                      time_scheme=scheme, shape=(n, n))
    a, b = tiled.step(h0, hb, 5), plain.step(h0, hb, 5)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    m = ~np.isnan(b)
    assert np.max(np.abs(a[m] - b[m])) <= 1e-15 * np.max(np.abs(b[m]))


@pytest.mark.parametrize("scheme", ["Euler", "rk2"])
def test_tiled_tangent_adjoint_pair(scheme):
    ny, nx = 44, 40
    h0, hb, ssh, rng = _fields(ny, nx, 2, land=True)
    mdt = 0.02 * rng.standard_normal((ny, nx))
    f = 1e-4 + 1e-6 * rng.standard_normal((ny, nx))
    model = TiledQgmOracle(10e3, 10e3, 1200.0, (ny, nx), c=2.5, f=f, SSH=ssh, mdt=mdt,
                           time_scheme=scheme, tile_size=20, tile_overlap=6)
    assert len(model.tiles) == 12  # 4 starts in y (the last repeated) x 3 in x
    land = np.isnan(ssh)
    dh = rng.standard_normal((ny, nx)) * 1e-3
    dh[land] = 0
    ad = rng.standard_normal((ny, nx))
    ad[land] = 0
    tl = model.step_tgl(dh, h0, hb, 4)
    aj = model.step_adj(ad, h0, hb, 4)
    lhs, rhs = np.nansum(tl * ad), np.nansum(dh * aj)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs))
    # centred finite differences of the forward agree with the tangent
    eps = 1e-6
    d = 50 * dh
    fd = (model.step(h0 + eps * d, hb, 4) - model.step(h0 - eps * d, hb, 4)) / (2 * eps)
    tl = model.step_tgl(d, h0, hb, 4)
    ok = ~np.isnan(fd)
    assert np.linalg.norm((fd - tl)[ok]) / np.linalg.norm(tl[ok]) < 1e-5


def test_plugin_module_exposes_tiled_class():
    """The reference instantiates getattr(module, config.MOD.name_class) of the
    jqgm.py it loads by path (mod.py:479-486); name_class may be 'QgmWithTiles'."""
    import os
    from importlib.machinery import SourceFileLoader
    import vardyn_amd
    path = os.path.join(os.path.dirname(vardyn_amd.__file__), "jqgm.py")
    mod = SourceFileLoader("jqgm_plugin_check", path).load_module()
    from vardyn_amd.tiles import QgmWithTiles
    assert getattr(mod, "QgmWithTiles") is QgmWithTiles
    assert hasattr(mod, "Qgm")
