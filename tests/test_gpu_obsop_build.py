"""Device construction of the L3 observation operator (vd_obs_grid_*,
vardyn_amd/obsop.py) against the oracle's scipy KDTree construction
(oracle/obsop.py sparse_op, obsop.py:417-460).

Bars: the selected rows and grid indices bit-exact, the distances bitwise
equal to the KDTree's (the same rounding of sqrt((dx^2 + dy^2) + dz^2)), the
weights within 1e-14 relative (the device exp against NumPy's).  Cases:
observations inside, on the edge of and far outside a regional grid,
Npix 2 / 4 / 8, a 128 x 128 grid with a dense swath;
an integer lattice with exact ties and points at exactly dmax against a
brute-force sort by (squared distance, grid index); Npix == 1 and empty
inputs give an empty operator and non-finite coordinates an error, as in the
reference.
"""
import numpy as np
import pytest
import torch

from oracle.obsop import geo2cart as ogeo2cart, sparse_op
from vardyn_amd import synthetic as syn
from vardyn_amd._lib import VardynError
from vardyn_amd.obsop import ObsGrid, geo2cart, interp_weights

pytestmark = pytest.mark.gpu


def _case(ny, nx, dx, n_obs, seed):
    g = syn.make_grid(ny, nx, dx)
    cc = geo2cart(np.column_stack((g.lon.ravel(), g.lat.ravel())))
    rng = np.random.default_rng(seed)
    iy, ix = rng.uniform(-3, ny + 2, n_obs), rng.uniform(-3, nx + 2, n_obs)
    lon, lat = g.index_to_lonlat(iy, ix)
    lon, lat = np.array(lon, dtype=float), np.array(lat, dtype=float)
    lon[:3] += 20.0  # far outside
    lon[4], lat[4] = g.lon.ravel()[17], g.lat.ravel()[17]  # on a grid point
    dmax = 4 * np.sqrt(2) * dx * 1e-3 * np.sqrt(2) / 2
    return cc, lon, lat, dmax


@pytest.mark.parametrize("npix", [2, 4, 8])
@pytest.mark.parametrize("shape", [(30, 26, 5e3, 400), (128, 128, 5e3, 20000)])
def test_tables_match_the_kdtree_oracle(npix, shape):
    cc, lon, lat, dmax = _case(*shape, seed=npix)
    grid = ObsGrid(cc, dmax, "cuda")
    w1, r1, c1 = interp_weights(grid, lon, lat, npix)
    w2, r2, c2 = sparse_op(cc, lon, lat, npix, dmax)
    assert w1.size > 0
    np.testing.assert_array_equal(r1, r2)
    np.testing.assert_array_equal(c1, c2)
    np.testing.assert_allclose(w1, w2, rtol=1e-14, atol=0)
    # the distances themselves, against the KDTree's
    from scipy.spatial import KDTree
    D, _ = KDTree(cc).query(ogeo2cart(np.column_stack((lon, lat))), k=npix)
    col, _, dist = grid.query(geo2cart(np.column_stack((lon, lat))), npix)
    torch.cuda.synchronize()
    col, dist = col.cpu().numpy(), dist.cpu().numpy()
    ok = col >= 0
    np.testing.assert_array_equal(ok, D <= dmax)
    np.testing.assert_array_equal(dist[ok], D[ok])
    assert np.all(np.isinf(dist[~ok]))


def test_lattice_ties_and_the_dmax_boundary():
    """Integer lattice in km: squared distances are exact integers, so exact
    ties are everywhere and points sit at exactly dmax = 2 (kept, D <= dmax);
    the order is (squared distance, grid index)."""
    g = np.stack(np.meshgrid(np.arange(9.0), np.arange(7.0), np.arange(3.0), indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(3)
    obs = np.concatenate([g[rng.integers(0, len(g), 40)], rng.integers(0, 9, (40, 3)) + 0.5,
                          rng.uniform(-1, 9, (40, 3))])
    dmax = 2.0
    grid = ObsGrid(g, dmax, "cuda")
    for npix in (1, 3, 8):
        col, w, dist = grid.query(obs, npix)
        torch.cuda.synchronize()
        col, w, dist = col.cpu().numpy(), w.cpu().numpy(), dist.cpu().numpy()
        for o, p in enumerate(obs):
            d = p - g
            d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
            keep = np.nonzero(np.sqrt(d2) <= dmax)[0]
            order = keep[np.lexsort((keep, d2[keep]))][:npix]
            n = order.size
            np.testing.assert_array_equal(col[o, :n], order)
            assert np.all(col[o, n:] == -1) and np.all(w[o, n:] == 0) and np.all(np.isinf(dist[o, n:]))
            np.testing.assert_array_equal(dist[o, :n], np.sqrt(d2[order]))
            if n:
                e = np.exp(-np.sqrt(d2[order]) ** 2 / (2 * (0.5 * dmax) ** 2))
                np.testing.assert_allclose(w[o, :n], e / e.sum(), rtol=1e-14)


def test_empty_and_rejected_inputs():
    cc, lon, lat, dmax = _case(30, 26, 5e3, 50, seed=0)
    grid = ObsGrid(cc, dmax, "cuda")
    assert interp_weights(grid, lon, lat, 1)[0].size == 0  # the reference's Npix == 1 quirk
    assert interp_weights(grid, lon[:0], lat[:0], 4)[0].size == 0
    far = interp_weights(grid, lon[:3], lat[:3], 4)
    assert far[0].size == 0
    with pytest.raises(VardynError):
        grid.query(geo2cart(np.column_stack((lon, lat))), 9)
    bad = lon.copy()
    bad[5] = np.nan  # the reference's KDTree query raises on it too
    with pytest.raises(VardynError):
        interp_weights(grid, bad, lat, 4)
    with pytest.raises(VardynError):
        ObsGrid(cc, float("nan"), "cuda")
