"""Pins of the FP64 restatement of Basis_wavelet3d (oracle/basis_wavelet3d.py).

The reference's own check is the dot-product test of operg against
operg_transpose (basis.py:4150-4158).  Beside it: the Gaspari-Cohn window's
closed form and integral, the isolated wavelet's closed-form field, the
structure of the centre lattice (spacing, symmetry about the domain's middle,
the first lattice point past each side outside the extended domain), the
direction count formula, the dense matrix built point by point on a tiny grid,
and the time-block layout of the coefficient vector.
"""
import numpy as np
import pytest
from scipy.integrate import quad

from oracle.basis_wavelet3d import KM2DEG, BasisWavelet3dOracle, gaspari_cohn, mywindow


def grid(ny=21, nx=25, lat0=40.0, dkm=8.0):
    lat1 = lat0 + (np.arange(ny) - (ny - 1) / 2) * dkm * KM2DEG
    lon1 = 3.0 + (np.arange(nx) - (nx - 1) / 2) * dkm * KM2DEG / np.cos(np.radians(lat0))
    return np.meshgrid(lon1, lat1)


KW = dict(lmin=40.0, lmax=170.0, npsp=3.5, facpsp=3.5, tmin=4.0, tmax=4.0)


def test_gaspari_cohn_closed_form():
    c = 3.7
    assert gaspari_cohn(0.0, c) == 1.0
    # r = c/2: both pieces give 5/24
    lo = -0.25 + 0.5 + 0.625 - 5 / 3 + 1
    hi = 1 / 12 - 0.5 + 0.625 + 5 / 3 - 5 + 4 - 2 / 3
    assert lo == pytest.approx(5 / 24, abs=1e-15) and hi == pytest.approx(5 / 24, abs=1e-15)
    assert gaspari_cohn(c / 2, c) == pytest.approx(5 / 24, abs=1e-15)
    assert gaspari_cohn(c, c) == pytest.approx(0.0, abs=1e-14)
    assert gaspari_cohn(1.01 * c, c) == 0.0 and gaspari_cohn(-5 * c, c) == 0.0
    # even, decreasing on [0, c]
    r = np.linspace(0, c, 200)
    g = gaspari_cohn(r, c)
    assert np.all(np.diff(g) <= 1e-15)
    assert np.array_equal(gaspari_cohn(-r, c), g)


def test_time_normalisation_is_the_window_integral():
    lon, lat = grid()
    b = BasisWavelet3dOracle(lon, lat, **KW)
    b.set_basis(np.arange(0, 9.0))
    for i, D in enumerate(b.DXt):
        exact = quad(lambda r: gaspari_cohn(r, D), -D, D, points=[-D / 2, 0, D / 2])[0]
        # closed form: 2 (D/2) [int_0^1 g + int_1^2 g] with the two polynomial pieces
        i1 = -1 / 24 + 1 / 10 + 5 / 32 - 5 / 9 + 1
        F = lambda x: x ** 6 / 72 - x ** 5 / 10 + 5 * x ** 4 / 32 + 5 * x ** 3 / 9 - 5 * x ** 2 / 2 + 4 * x \
            - 2 / 3 * np.log(x)
        assert exact == pytest.approx(D * (i1 + F(2.0) - F(1.0)), rel=1e-10)
        # the reference integrates with a left Riemann sum over 50 points
        assert b.norm_fact[i] == pytest.approx(exact, rel=2e-3)


def test_ladder_and_directions():
    lon, lat = grid()
    b = BasisWavelet3dOracle(lon, lat, **KW)
    b.set_basis(np.arange(0, 9.0))
    # ratio 1 + facpsp/npsp = 2: wavelengths 40, 80, 160 (320 > lmax dropped)
    np.testing.assert_allclose(1 / b.fs, [160.0, 80.0, 40.0], rtol=1e-12)
    # linspace(0, pi, int(pi f0/df0 facpsp)) without its last value: f0/df0 = 1
    assert b.ntheta == int(np.pi * 3.5) - 1
    np.testing.assert_allclose(b.theta, np.arange(b.ntheta) * np.pi / b.ntheta, atol=1e-15)
    np.testing.assert_allclose(b.DXs, 1 / b.fs * 3.5 / 2)


def test_centre_lattice_structure():
    lon, lat = grid()
    b = BasisWavelet3dOracle(lon, lat, **KW)
    b.set_basis(np.arange(0, 9.0))
    mid_lat = (lat.min() + lat.max()) / 2
    mid_lon = (lon.min() + lon.max()) / 2
    for f in range(b.nfs):
        D = b.DXs[f]
        rows = np.unique(b.ENSLAT[f])
        step = D * KM2DEG
        np.testing.assert_allclose(np.diff(rows), step, rtol=1e-9)
        assert np.min(np.abs(rows - mid_lat)) < 1e-12  # the middle is a row
        # the next row past each end would lie beyond the extended domain
        assert rows[0] - step <= lat.min() - D * KM2DEG + 1e-12
        assert rows[-1] + step >= lat.max() + D * KM2DEG - 1e-12
        assert rows[0] > lat.min() - D * KM2DEG and rows[-1] < lat.max() + D * KM2DEG
        for la in rows:
            lons = b.ENSLON[f][b.ENSLAT[f] == la]
            s = D * KM2DEG / np.cos(np.radians(la))
            np.testing.assert_allclose(np.diff(lons), s, rtol=1e-9)
            assert np.min(np.abs(lons - mid_lon)) < 1e-12
    # NT: time centres DXt apart around the middle of the span
    t = np.arange(0, 9.0)
    for i, e in enumerate(b.ENST):
        np.testing.assert_allclose(np.diff(e), b.DXt[i], rtol=1e-12)
        assert np.min(np.abs(e - 4.0)) < 1e-12
        assert e[0] > t.min() - b.DXt[i] and e[-1] < t.max() + b.DXt[i]


def test_dense_matrix_point_by_point():
    lon, lat = grid(ny=9, nx=11)
    b = BasisWavelet3dOracle(lon, lat, lmin=60.0, lmax=130.0, npsp=3.5, facpsp=3.5, tmin=4.0, tmax=4.0)
    b.set_basis(np.arange(0, 5.0))
    lon1, lat1 = lon.ravel(), lat.ravel()
    for f in range(b.nfs):
        dense = b.Gx[f].toarray()
        col = 0
        for P in range(b.NS[f]):
            lo, la = b.ENSLON[f][P], b.ENSLAT[f][P]
            cl = np.cos(la * np.pi / 180)
            for th in b.theta:
                for trig in (np.cos, np.sin):
                    ref = np.zeros(lon1.size)
                    for p in range(lon1.size):
                        dx = (lon1[p] - lo) / KM2DEG * cl
                        dy = (lat1[p] - la) / KM2DEG
                        if abs(dx) <= b.DXs[f] and abs(dy) <= b.DXs[f]:
                            w = np.cos(np.pi / 2 * dx / b.DXs[f]) ** 2 * np.cos(np.pi / 2 * dy / b.DXs[f]) ** 2
                            ref[p] = np.sqrt(2) * w * trig(b.k[f] * (dx * np.cos(th) + dy * np.sin(th)))
                    np.testing.assert_allclose(dense[:, col], ref, rtol=1e-12, atol=1e-14)
                    col += 1


def test_isolated_wavelet_closed_form():
    lon, lat = grid(ny=41, nx=41)
    b = BasisWavelet3dOracle(lon, lat, **KW)
    times = np.arange(0, 9.0)
    b.set_basis(times)
    f = b.nfs - 1  # the shortest waves: many centres, one sits on the grid's middle point
    mid = np.argmin(np.abs(b.ENSLON[f] - lon[20, 20]) + np.abs(b.ENSLAT[f] - lat[20, 20]))
    assert abs(b.ENSLON[f][mid] - lon[20, 20]) < 1e-12 and abs(b.ENSLAT[f][mid] - lat[20, 20]) < 1e-12
    q2 = 2 * b.ntheta
    nx_f = q2 * b.NS[f]
    # time block 1 (the middle centre of the one time frequency), direction 0
    ib = 1
    it = 4  # t = 4 days: dt = 0 for the middle time centre
    assert b.ENST[0][ib] == pytest.approx(4.0)
    for cs, val in ((0, np.sqrt(2)), (1, 0.0)):
        X = np.zeros(b.nbasis)
        X[b.fs_wavebounds[f] + ib * nx_f + mid * q2 + cs] = 1.0
        phi = b.operg(it, X)
        g = 1.0 / b.norm_fact[0]
        assert phi[20, 20] == pytest.approx(val * g, abs=1e-14)
        # one point east along theta = 0: sqrt(2) w(dx/D) cos(k dx) (sin for cs = 1)
        dx = (lon[20, 21] - lon[20, 20]) / KM2DEG * np.cos(np.radians(lat[20, 20]))
        expect = np.sqrt(2) * mywindow(dx / b.DXs[f]) * (np.cos if cs == 0 else np.sin)(b.k[f] * dx) * g
        assert phi[20, 21] == pytest.approx(expect, rel=1e-12, abs=1e-15)
        # outside the box: zero
        far = np.abs((lon - lon[20, 20]) / KM2DEG * np.cos(np.radians(lat[20, 20]))) > b.DXs[f]
        assert np.all(phi[far] == 0.0)


@pytest.mark.parametrize("masked", [False, True])
def test_reference_dot_product(masked):
    lon, lat = grid()
    rng = np.random.default_rng(7)
    mask = (rng.random(lon.shape) < 0.2) if masked else None
    b = BasisWavelet3dOracle(lon, lat, mask=mask, **KW)
    times = np.arange(0, 9.0)
    b.set_basis(times)
    for it in (0, 3, 8):
        psi = rng.random(b.nbasis)
        phi = rng.random(lon.shape)
        ps1 = np.inner(psi, b.operg_transpose(it, phi))
        ps2 = np.inner(b.operg(it, psi).ravel(), phi.ravel())
        assert ps1 == pytest.approx(ps2, rel=1e-12)
    if masked:
        assert np.all(b.operg(2, rng.random(b.nbasis))[mask] == 0.0)


def test_time_blocks_layout():
    """Coefficient block b of frequency f at time t is weighted by
    GC(t - t_b, DXt)/norm: a constant coefficient vector over one frequency
    gives that frequency's space sum times the sum of the active factors."""
    lon, lat = grid()
    b = BasisWavelet3dOracle(lon, lat, **KW)
    times = np.arange(0, 9.0)
    b.set_basis(times)
    f = 0
    q2 = 2 * b.ntheta
    nx_f = q2 * b.NS[f]
    v = np.random.default_rng(3).standard_normal(nx_f)
    X = np.zeros(b.nbasis)
    nb = sum(e.size for e in b.ENST)
    for ib in range(nb):
        X[b.fs_wavebounds[f] + ib * nx_f: b.fs_wavebounds[f] + (ib + 1) * nx_f] = v
    space = b.Gx[f].dot(v).reshape(lon.shape)
    for it, t in enumerate(times):
        gsum = sum(gaspari_cohn(t - tc, b.DXt[0]) / b.norm_fact[0] for tc in b.ENST[0] if abs(t - tc) < b.DXt[0])
        np.testing.assert_allclose(b.operg(it, X), gsum * space, rtol=1e-12, atol=1e-14)
