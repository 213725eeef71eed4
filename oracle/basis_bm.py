"""FP64 restatement of the balanced-motion wavelet basis ``Basis_bm``.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``mapping/src/basis.py`` of datlas-ocean/VarDyn:

* ``mywindow`` / ``mywindow_flux`` -- basis.py:4138-4144: cos^2(pi x / 2) and
  -pi sin(pi x / 2) cos(pi x / 2) on [-1, 1].
* ``BasisBmOracle.set_basis`` -- basis.py:564-724: the pseudo-frequencies
  (log-spaced from 1/lmin down to 1/lmax, ascending), the directions, and for
  every frequency the wavelet centres on latitude rows spaced DX/facns (km)
  with per-row longitude spacing, dropping centres whose (1/f)-box holds only
  masked points (:640-655); decorrelation times (:658-664), time centres
  (:666-667), the time-window integral used for normalisation (:669-673), and
  the prior variances Q (:679-697).
* ``_space`` -- basis.py:726-795: for every centre the grid points with
  |dx| <= DX and |dy| <= DX (km, dx scaled by the cosine of the centre's
  latitude), minus masked points, tapered by mywindow(dx/DX) mywindow(dy/DX),
  and for every direction a cosine and a sine column sqrt(2) taper
  cos/sin(k . (dx, dy)); the columns of one frequency form a CSC matrix of
  shape (grid points, 2 ntheta NP) in the order (centre, direction, cos/sin).
* ``_time`` -- basis.py:797-820: per time t and frequency, the factor
  window((t - t_c) / tdec) (/ norm_fact with ``norm_time``) of every time
  centre with |t - t_c| < tdec and a non-zero factor, repeated over the
  frequency's 2 ntheta NP space columns.
* ``operg`` / ``operg_transpose`` -- basis.py:863-936 (SSH only): the field
  sum over frequencies of Gx_f (sum over active time centres of
  factor x coefficient block), and its transpose.

* ``geostrophic_tables`` / ``ssh2uv`` / ``ssh2uv_adj`` -- basis.py:514-526,
  822-860 (``compute_velocities``): f, dx, dy padded by one edge row and column
  and averaged onto the u and v grids; u = -g/f_u dP_y / dy_u and
  v = g/f_v dP_x / dx_v of the edge-padded SSH P, and the transpose with the
  padded row, column and corner folded back onto the SSH's edge.

Pinning (tests/test_oracle_basis_bm.py): the reference's own dot-product
test of operg against operg_transpose (basis.py:4150-4158); closed forms of
the windows, of one isolated wavelet's field and of the time factors; the
counts of frequencies, directions and coefficients from the formulas; the Q
profile (constant above lmeso, power law below); a dense brute-force
construction of Gx on a tiny grid equal to the CSC one; for the velocities,
the dot-product test of ssh2uv against ssh2uv_adj, the exact velocities of
a tilted plane on a grid with dx != dy and f varying along y (interior
values, zero first row of u and first column of v), and the midpoint tables
of a linear f.
"""
from __future__ import annotations

import numpy as np
from scipy.sparse import csc_matrix

KM2DEG = 1.0 / 110


def mywindow(x):
    """basis.py:4138-4140."""
    return np.cos(x * 0.5 * np.pi) ** 2


def mywindow_flux(x):
    """basis.py:4142-4144."""
    return -np.pi * np.sin(x * 0.5 * np.pi) * np.cos(x * 0.5 * np.pi)


class BasisBmOracle:
    """``Basis_bm`` on a grid given by 2-D ``lon``, ``lat`` (degrees) and an
    optional boolean ``mask`` (True = masked), SSH component only."""

    controls_initial_state = False
    time_dependent = True  # operg(t_days, X) / operg_transpose(t_days, ad)

    def __init__(self, lon, lat, mask=None, flux=False, facns=1.0, facnlt=2.0, npsp=3.5,
                 facpsp=1.5, lmin=80.0, lmax=970.0, lmeso=300.0, tmeso=20.0, sloptdec=-1.28,
                 factdec=0.5, tdecmin=2.5, tdecmax=40.0, Qmax=1e-3, slopQ=-5.0, norm_time=True,
                 delta_time_ref=0.0):
        self.lon1d = np.asarray(lon, dtype=np.float64).ravel()
        self.lat1d = np.asarray(lat, dtype=np.float64).ravel()
        self.shape_phys = np.shape(lon)
        self.nphys = self.lon1d.size
        self.lon_min, self.lon_max = self.lon1d.min(), self.lon1d.max()
        self.lat_min, self.lat_max = self.lat1d.min(), self.lat1d.max()
        m = None if mask is None else np.asarray(mask, dtype=bool).ravel()
        self.mask1d = m if (m is not None and np.any(m)) else None
        self.flux, self.facns, self.facnlt, self.npsp, self.facpsp = flux, facns, facnlt, npsp, facpsp
        self.lmin, self.lmax, self.lmeso, self.tmeso = lmin, lmax, lmeso, tmeso
        self.sloptdec, self.factdec, self.tdecmin, self.tdecmax = sloptdec, factdec, tdecmin, tdecmax
        self.Qmax, self.slopQ, self.norm_time = Qmax, slopQ, norm_time
        self.delta_time_ref = delta_time_ref
        self.window = mywindow_flux if flux else mywindow

    def set_basis(self, time, return_q=False):
        time = np.asarray(time, dtype=np.float64)
        LON_MIN, LON_MAX, LAT_MIN, LAT_MAX = self.lon_min, self.lon_max, self.lat_min, self.lat_max
        if LON_MAX < LON_MIN:
            LON_MAX = LON_MAX + 360.0
        step = np.log(1 + self.facpsp / self.npsp)
        logff = np.arange(np.log(1.0 / self.lmin), np.log(1.0 / self.lmax) - step, -step)[::-1]
        ff = np.exp(logff)
        dff = ff[1:] - ff[:-1]
        theta = np.linspace(0, np.pi, int(np.pi * ff[0] / dff[0] * self.facpsp))[:-1]
        nf, ntheta = len(ff), len(theta)
        deltat = time.max() - time.min()
        DX = 1.0 / ff * self.npsp * 0.5
        DXG = DX / self.facns
        ENSLON, ENSLAT, enst, tdec, norm_fact = [], [], [], [], []
        NP = np.empty(nf, dtype=np.int64)
        for iff in range(nf):
            lons, lats = [], []
            dlat = DXG[iff] * KM2DEG
            lat0 = LAT_MIN - LAT_MIN % dlat - DX[iff] * KM2DEG
            lat1 = LAT_MAX + DX[iff] * KM2DEG * 1.5
            for la in np.arange(lat0, lat1, dlat):
                c = np.cos(la * np.pi / 180.0)
                dlon = DXG[iff] / c * KM2DEG
                lon0 = LON_MIN - LON_MIN % dlon - DX[iff] / c * KM2DEG
                lon1 = LON_MAX + DX[iff] / c * KM2DEG * 1.5
                for lo in np.arange(lon0, lon1, dlon):
                    if self.mask1d is not None:
                        ind = np.where(
                            (np.abs((self.lon1d - lo) / KM2DEG * np.cos(la * np.pi / 180.0)) <= 1.0 / ff[iff]) &
                            (np.abs((self.lat1d - la) / KM2DEG) <= 1.0 / ff[iff]))[0]
                        if np.all(self.mask1d[ind]):
                            continue
                    lons.append(lo)
                    lats.append(la)
            ENSLON.append(np.array(lons))
            ENSLAT.append(np.array(lats))
            NP[iff] = len(lons)
            td = self.tmeso * self.lmeso ** self.sloptdec * ff[iff] ** self.sloptdec * self.factdec
            td = min(max(td, self.tdecmin), self.tdecmax)
            tdec.append(td)
            t0 = -self.delta_time_ref % td
            enst.append(np.arange(t0 - td / self.facnlt, deltat + td / self.facnlt, td / self.facnlt))
            tt = np.linspace(-td, td)
            acc = np.zeros_like(tt)
            for i in range(tt.size - 1):
                acc[i + 1] = acc[i] + self.window(tt[i] / td) * (tt[i + 1] - tt[i])
            norm_fact.append(acc.max())
        Q, bounds, iwave = [], [], 0
        for iff in range(nf):
            bounds.append(iwave)
            if NP[iff] > 0:
                n = 2 * len(enst[iff]) * ntheta * NP[iff]
                if 1.0 / ff[iff] > self.lmeso:
                    Q.append(self.Qmax / self.facns * np.ones(n))
                else:
                    Q.append(self.Qmax / self.facns * self.lmeso ** self.slopQ * ff[iff] ** self.slopQ * np.ones(n))
                iwave += n
        bounds.append(iwave)
        self.ff, self.theta, self.nf, self.ntheta, self.k = ff, theta, nf, ntheta, 2 * np.pi * ff
        self.DX, self.ENSLON, self.ENSLAT, self.NP = DX, ENSLON, ENSLAT, NP
        self.tdec, self.enst, self.norm_fact = tdec, enst, norm_fact
        self.iff_wavebounds = bounds
        self.Q = np.concatenate(Q) if Q else np.zeros(0)
        self.nbasis = self.Q.size
        self.Gx, self.Nx = self._space()
        self.Gt, self.Nt = self._time(time)
        if return_q:
            return np.zeros_like(self.Q), self.Q

    def _space(self):
        Gx, Nx = [], []
        for iff in range(self.nf):
            data, indices, sizes = [], [], []
            for P in range(self.NP[iff]):
                lo, la = self.ENSLON[iff][P], self.ENSLAT[iff][P]
                ind = np.where(
                    (np.abs((self.lon1d - lo) / KM2DEG * np.cos(la * np.pi / 180.0)) <= self.DX[iff]) &
                    (np.abs((self.lat1d - la) / KM2DEG) <= self.DX[iff]))[0]
                xx = (self.lon1d[ind] - lo) / KM2DEG * np.cos(la * np.pi / 180.0)
                yy = (self.lat1d[ind] - la) / KM2DEG
                if self.mask1d is not None:
                    keep = ~self.mask1d[ind]
                    ind, xx, yy = ind[keep], xx[keep], yy[keep]
                facs = mywindow(xx / self.DX[iff]) * mywindow(yy / self.DX[iff])
                for th in self.theta:
                    kx = self.k[iff] * np.cos(th)
                    ky = self.k[iff] * np.sin(th)
                    for fn in (np.cos, np.sin):
                        sizes.append(ind.size)
                        indices.append(ind)
                        data.append(np.sqrt(2) * facs * fn(kx * xx + ky * yy))
            n = len(sizes)
            Nx.append(n)
            indptr = np.zeros(n + 1, dtype=np.int64)
            indptr[1:] = np.cumsum(sizes)
            Gx.append(csc_matrix((np.concatenate(data) if data else np.zeros(0),
                                  np.concatenate(indices) if indices else np.zeros(0, dtype=np.int64),
                                  indptr), shape=(self.nphys, n)))
        return Gx, Nx

    def _time(self, time):
        Gt, Nt = {}, {}
        for t in time:
            Gt[t], Nt[t] = [], []
            for iff in range(self.nf):
                blk = 2 * self.ntheta * self.NP[iff]
                g = np.zeros(self.iff_wavebounds[iff + 1] - self.iff_wavebounds[iff])
                n = 0
                for it, tc in enumerate(self.enst[iff]):
                    dt = t - tc
                    if abs(dt) < self.tdec[iff]:
                        fact = self.window(dt / self.tdec[iff])
                        if self.norm_time:
                            fact /= self.norm_fact[iff]
                        if fact != 0:
                            n += 1
                            g[it * blk:(it + 1) * blk] = fact
                Gt[t].append(g)
                Nt[t].append(n)
        return Gt, Nt

    def operg(self, t, X):
        X = np.asarray(X, dtype=np.float64)
        ssh = np.zeros(self.nphys)
        for iff in range(self.nf):
            Xf = X[self.iff_wavebounds[iff]:self.iff_wavebounds[iff + 1]]
            g = self.Gt[t][iff]
            ind0 = np.nonzero(g)[0]
            if ind0.size > 0:
                blocks = (g * Xf)[ind0].reshape(self.Nt[t][iff], self.Nx[iff])
                ssh += self.Gx[iff].dot(blocks.sum(axis=0))
        return ssh.reshape(self.shape_phys)

    def operg_transpose(self, t, adssh):
        adssh = np.asarray(adssh, dtype=np.float64).ravel()
        adX = np.zeros(self.nbasis)
        for iff in range(self.nf):
            g = self.Gt[t][iff]
            ind0 = np.nonzero(g)[0]
            if ind0.size > 0:
                gb = g[ind0].reshape(self.Nt[t][iff], self.Nx[iff])
                col = self.Gx[iff].T.dot(adssh)
                sl = slice(self.iff_wavebounds[iff], self.iff_wavebounds[iff + 1])
                adX[sl][ind0] += (gb * col[None, :]).ravel()
        return adX


def geostrophic_tables(f, DX, DY, g=9.81):
    """basis.py:514-526: (au, dy_on_u, bv, dx_on_v) with au = -g / f_on_u and
    bv = g / f_on_v; f, DX, DY (ny, nx) padded by one edge row and column."""
    pad = ((1, 0), (1, 0))
    _f = np.pad(np.asarray(f, dtype=np.float64), pad_width=pad, mode="edge")
    f_on_v = 0.5 * (_f[:, 1:] + _f[:, :-1])
    f_on_u = 0.5 * (_f[1:, :] + _f[:-1, :])
    dx = np.pad(np.asarray(DX, dtype=np.float64), pad_width=pad, mode="edge")
    dy = np.pad(np.asarray(DY, dtype=np.float64), pad_width=pad, mode="edge")
    dx_on_v = 0.5 * (dx[:, 1:] + dx[:, :-1])
    dy_on_u = 0.5 * (dy[1:, :] + dy[:-1, :])
    return -g / f_on_u, dy_on_u, g / f_on_v, dx_on_v


def ssh2uv(ssh, au, dy_on_u, bv, dx_on_v):
    """basis.py:822-833: u (ny, nx+1), v (ny+1, nx) of one SSH field."""
    _ssh = np.pad(np.asarray(ssh, dtype=np.float64), pad_width=((1, 0), (1, 0)), mode="edge")
    u = au * (_ssh[1:, :] - _ssh[:-1, :]) / dy_on_u
    v = bv * (_ssh[:, 1:] - _ssh[:, :-1]) / dx_on_v
    return u, v


def ssh2uv_adj(adu, adv, au, dy_on_u, bv, dx_on_v):
    """basis.py:835-860: the SSH adjoint of (adu, adv)."""
    ny, nx1 = np.shape(adu)
    _adssh = np.zeros((ny + 1, nx1))
    _adssh[1:, :] += au * adu / dy_on_u
    _adssh[:-1, :] += -au * adu / dy_on_u
    _adssh[:, 1:] += bv * adv / dx_on_v
    _adssh[:, :-1] += -bv * adv / dx_on_v
    adssh = _adssh[1:, 1:].copy()
    adssh[0, :] += _adssh[0, 1:]
    adssh[:, 0] += _adssh[1:, 0]
    adssh[0, 0] += _adssh[0, 0]
    return adssh
