"""Pins of the FP64 ``Basis_bm`` oracle (mapping/src/basis.py:472-936,
4138-4158) before it is trusted as the GPU basis's checker: the reference's
own dot-product test of operg against operg_transpose, closed forms of the
windows and of the time structure, the counts of frequencies, directions and
coefficients, the Q profile, and a brute-force dense construction of the
space matrix on a tiny grid."""
import numpy as np
import pytest

from oracle.basis_bm import KM2DEG, BasisBmOracle, mywindow, mywindow_flux


def small_grid(ny=28, nx=32, d=0.05, land=False):
    lon, lat = np.meshgrid(-60 + d * np.arange(nx), 35 + d * np.arange(ny))
    mask = None
    if land:
        mask = np.zeros((ny, nx), dtype=bool)
        mask[:6, :7] = True
    return lon, lat, mask


def small_basis(flux=False, land=False, times=np.arange(0.0, 6.0)):
    lon, lat, mask = small_grid(land=land)
    b = BasisBmOracle(lon, lat, mask, flux=flux, lmin=30.0, lmax=110.0, delta_time_ref=3.3)
    b.set_basis(times)
    return b


def test_windows_closed_forms():
    x = np.linspace(-1, 1, 101)
    assert np.allclose(mywindow(x), 0.5 * (1 + np.cos(np.pi * x)), atol=1e-15)
    assert np.allclose(mywindow_flux(x), -0.5 * np.pi * np.sin(np.pi * x), atol=1e-15)
    assert mywindow(0.0) == 1.0 and abs(mywindow(1.0)) < 1e-30


def test_frequencies_directions_and_counts():
    b = small_basis()
    ratio = 1 + 1.5 / 3.5
    assert np.allclose(b.ff[1:] / b.ff[:-1], ratio, rtol=1e-12)
    assert abs(b.ff[-1] - 1 / 30.0) < 1e-15          # the shortest wavelength is lmin
    assert 1 / b.ff[0] <= 110.0 * ratio and 1 / b.ff[0] > 110.0 / ratio
    # directions: [0, pi), as many as pi f0 / df0 * facpsp (minus the end point)
    n = int(np.pi * b.ff[0] / (b.ff[1] - b.ff[0]) * 1.5)
    assert b.ntheta == n - 1 and np.allclose(b.theta, np.arange(n - 1) * np.pi / (n - 1))
    nb = sum(2 * b.ntheta * b.NP[f] * len(b.enst[f]) for f in range(b.nf))
    assert b.nbasis == nb == b.iff_wavebounds[-1]
    for f in range(b.nf):
        assert b.Gx[f].shape == (b.nphys, 2 * b.ntheta * b.NP[f])
        assert np.allclose(np.diff(b.enst[f]), b.tdec[f] / 2.0)   # facnlt = 2
        # wavelet centres tile the domain: spacing DX / facns in latitude
        lats = np.unique(b.ENSLAT[f])
        assert np.allclose(np.diff(lats), b.DX[f] * KM2DEG)


def test_q_profile_and_decorrelation_times():
    b = small_basis()
    for f in range(b.nf):
        q = b.Q[b.iff_wavebounds[f]:b.iff_wavebounds[f + 1]]
        lam = 1 / b.ff[f]
        want = 1e-3 if lam > 300.0 else 1e-3 * (lam / 300.0) ** 5   # Qmax (lmeso f)^slopQ
        assert np.allclose(q, want, rtol=1e-12)
        td = 20 * 0.5 * (lam / 300.0) ** 1.28
        assert b.tdec[f] == pytest.approx(min(max(td, 2.5), 40.0), rel=1e-12)


@pytest.mark.parametrize("flux", [False, True])
def test_time_normalisation(flux):
    """The normalisation is the window's integral over [-tdec, tdec] (left
    Riemann sum of 50 points): 1 x tdec for cos^2, and the flux window's
    cumulative maximum, also 1 x tdec (integral of -pi/2 sin(pi x) over [-1, 0])."""
    b = small_basis(flux=flux)
    for f in range(b.nf):
        assert b.norm_fact[f] == pytest.approx(b.tdec[f], rel=0.05)


@pytest.mark.parametrize("flux,land", [(False, False), (True, False), (False, True)])
def test_operg_adjoint_identity(flux, land):
    """basis.py:4150-4158: <operg(t, psi), phi> = <psi, operg^T(t, phi)>."""
    b = small_basis(flux=flux, land=land)
    rng = np.random.default_rng(0)
    for t in (0.0, 2.0, 5.0):
        psi = rng.standard_normal(b.nbasis)
        phi = rng.standard_normal(b.shape_phys)
        lhs = np.sum(b.operg(t, psi) * phi)
        rhs = np.dot(psi, b.operg_transpose(t, phi))
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_space_matrix_brute_force():
    """Every column of Gx against a point-by-point evaluation of the wavelet
    (sqrt 2 w(dx/DX) w(dy/DX) cos|sin(k (dx cos th + dy sin th)) on the box,
    0 elsewhere and on masked points), written as explicit loops."""
    lon, lat, mask = small_grid(ny=14, nx=16, land=True)
    b = BasisBmOracle(lon, lat, mask, lmin=30.0, lmax=60.0)
    b.set_basis(np.arange(0.0, 3.0))
    for f in range(b.nf):
        G = b.Gx[f].toarray()
        col = 0
        for P in range(b.NP[f]):
            lo, la = b.ENSLON[f][P], b.ENSLAT[f][P]
            for th in b.theta:
                for kind in (0, 1):
                    want = np.zeros(b.nphys)
                    for p in range(b.nphys):
                        i, j = divmod(p, lon.shape[1])
                        dx = (lon[i, j] - lo) * 110.0 * np.cos(np.radians(la))
                        dy = (lat[i, j] - la) * 110.0
                        if abs(dx) > b.DX[f] or abs(dy) > b.DX[f] or mask[i, j]:
                            continue
                        w = np.cos(np.pi / 2 * dx / b.DX[f]) ** 2 * np.cos(np.pi / 2 * dy / b.DX[f]) ** 2
                        ph = b.k[f] * (dx * np.cos(th) + dy * np.sin(th))
                        want[p] = np.sqrt(2) * w * (np.cos(ph) if kind == 0 else np.sin(ph))
                    assert np.allclose(G[:, col], want, rtol=1e-9, atol=1e-12)
                    col += 1


def test_single_wavelet_field_and_time_factor():
    """One unit coefficient at its time centre: the field is the wavelet
    divided by the time normalisation; half a decorrelation time away the
    factor is cos^2(pi/4) = 1/2 of that; at or beyond tdec it is zero."""
    b = small_basis(times=np.arange(0.0, 12.0, 0.25))
    f = b.nf - 1
    it = 2
    tc = b.enst[f][it]
    X = np.zeros(b.nbasis)
    j = b.iff_wavebounds[f] + it * 2 * b.ntheta * b.NP[f]      # centre 0, direction 0, cosine
    X[j] = 1.0
    wave = b.Gx[f][:, 0].toarray().ravel()
    for dt, fac in ((0.0, 1.0), (0.5, 0.5), (1.0, 0.0)):
        t = tc + dt * b.tdec[f]
        if t not in b.Gt:
            b.Gt.update(b._time(np.array([t]))[0])
            b.Nt.update(b._time(np.array([t]))[1])
        got = b.operg(t, X).ravel()
        assert np.allclose(got, fac * wave / b.norm_fact[f], rtol=1e-12, atol=1e-15)
