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


# ------------------------------------------------- geostrophic velocities (basis.py:514-526, 822-860)
def _uv_grid(ny=9, nx=12, seed=3):
    from oracle.basis_bm import geostrophic_tables
    rng = np.random.default_rng(seed)
    f = 1e-4 + 2e-6 * rng.standard_normal((ny, nx))
    DX = 5e3 * (1 + 0.1 * rng.random((ny, nx)))
    DY = 7e3 * (1 + 0.1 * rng.random((ny, nx)))
    return rng, geostrophic_tables(f, DX, DY)


def test_ssh2uv_dot_product():
    from oracle.basis_bm import ssh2uv, ssh2uv_adj
    rng, tab = _uv_grid()
    ssh = rng.standard_normal((9, 12))
    adu, adv = rng.standard_normal((9, 13)), rng.standard_normal((10, 12))
    u, v = ssh2uv(ssh, *tab)
    lhs = np.sum(u * adu) + np.sum(v * adv)
    rhs = np.sum(ssh * ssh2uv_adj(adu, adv, *tab))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def test_ssh2uv_tilted_plane_closed_form():
    """ssh = a x + b y on a grid with dx != dy and f linear in y: u = -g b / f_u
    and v = g a / f_v away from the padded edge, u = 0 on the first row and
    v = 0 on the first column (the edge padding repeats the first row and
    column); f_u is the midpoint of consecutive rows of f, f_v the row's f."""
    from oracle.basis_bm import geostrophic_tables, ssh2uv
    ny, nx, dx, dy, g = 8, 11, 5e3, 7e3, 9.81
    fr = 1e-4 + 3e-6 * np.arange(ny)
    f = fr[:, None] * np.ones((1, nx))
    tab = geostrophic_tables(f, dx * np.ones((ny, nx)), dy * np.ones((ny, nx)))
    a, b = 2e-6, -3e-6
    X, Y = np.meshgrid(dx * np.arange(nx), dy * np.arange(ny))
    u, v = ssh2uv(a * X + b * Y, *tab)
    fu = 0.5 * (fr + np.concatenate([fr[:1], fr[:-1]]))  # f_on_u row r: mean of f rows r and r-1
    np.testing.assert_allclose(u[1:], (-g * b / fu[1:])[:, None] * np.ones((1, nx + 1)), rtol=1e-12)
    assert np.all(u[0] == 0.0)
    fv = np.concatenate([fr[:1], fr])  # f_on_v row r: f row max(r-1, 0)
    np.testing.assert_allclose(v[:, 1:], (g * a / fv)[:, None] * np.ones((1, nx - 1)), rtol=1e-12)
    assert np.all(v[:, 0] == 0.0)


def test_geostrophic_tables_midpoints():
    from oracle.basis_bm import geostrophic_tables
    ny, nx = 5, 6
    f = 1e-4 * (1 + 0.1 * np.arange(ny))[:, None] * np.ones((1, nx))
    DX = np.ones((ny, nx)) * (1000.0 + 10 * np.arange(nx))[None, :]
    DY = np.ones((ny, nx)) * (2000.0 + 20 * np.arange(ny))[:, None]
    au, dyu, bv, dxv = geostrophic_tables(f, DX, DY)
    assert au.shape == dyu.shape == (ny, nx + 1) and bv.shape == dxv.shape == (ny + 1, nx)
    np.testing.assert_allclose(-9.81 / au[2, 3], 0.5 * (f[2, 0] + f[1, 0]), rtol=1e-15)
    np.testing.assert_allclose(dyu[0, 0], DY[0, 0], rtol=1e-15)
    np.testing.assert_allclose(dyu[3, 2], 0.5 * (DY[3, 0] + DY[2, 0]), rtol=1e-15)
    np.testing.assert_allclose(dxv[2, 0], DX[0, 0], rtol=1e-15)
    np.testing.assert_allclose(dxv[2, 4], 0.5 * (DX[0, 4] + DX[0, 3]), rtol=1e-15)
