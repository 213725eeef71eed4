"""FP64 restatement of the 3-D wavelet basis ``Basis_wavelet3d``.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``mapping/src/basis.py`` of datlas-ocean/VarDyn:

* ``mywindow`` -- basis.py:4138-4140: cos^2(pi x / 2) on [-1, 1].
* ``gaspari_cohn`` -- mapping/src/tools.py:16-47: the fifth-order piecewise
  rational function of 2|r|/c, zero beyond |r| = c.
* ``BasisWavelet3dOracle.set_basis`` -- basis.py:2681-2848: space
  pseudo-frequencies (log-spaced from 1/lmin down to 1/lmax with ratio
  1 + facpsp/npsp, ascending, wavelengths above lmax dropped), time
  pseudo-frequencies (the same ladder between 1/tmin and 1/tmax), the
  directions (linspace(0, pi, int(pi f_0/df_0 facpsp)) without its last
  value), for every space frequency the centres on latitude rows DX_f/facnst
  (km) apart around the domain's middle, longitudes DX_f/facnst/cos(lat)
  apart on each row, centres whose (1/(2f))-box holds only masked points
  dropped when there is a mask (:2749-2768); for every time frequency the
  time centres DXt/facnst apart around the middle of the time span and the
  left Riemann sum of the Gaspari-Cohn window over 50 points used as its
  normalisation (:2770-2791); Q = sigma_Q for every coefficient (:2793-2806);
  the coefficients of frequency f laid out as (time frequency, time centre,
  space centre, direction, cos/sin).
* ``_space`` -- basis.py:2850-2910: for every centre the grid points with
  |dx| <= DX_f and |dy| <= DX_f (km, dx scaled by the cosine of the centre's
  latitude), minus masked points, tapered by mywindow(dx/DX_f)
  mywindow(dy/DX_f); for every direction a cosine and a sine column
  sqrt(2) taper cos/sin(k (dx cos theta + dy sin theta)), k = 2 pi f; the
  columns of one frequency form a CSC matrix (grid points, 2 ntheta NS_f).
* ``_time`` -- basis.py:2912-2934: per basis time t and space frequency, the
  factor GC(t - t_c, DXt)/norm of every time centre with |t - t_c| < DXt and
  a non-zero factor, over that centre's block of the frequency's space
  columns.
* ``operg`` / ``operg_transpose`` -- basis.py:2936-2985: the field sum over
  frequencies of Gx_f (sum over the active time blocks of factor x
  coefficient block), and its transpose.  The reference keys the time
  factors by the float value of t; here a basis time is given by its index
  in the ``time`` array of ``set_basis``.

Pinning (tests/test_oracle_basis_wavelet3d.py): the reference's dot-product
test of operg against operg_transpose (basis.py:4150-4158); a dense
brute-force construction of every space column from the formula on a tiny
grid equal to the CSC matrix; the Gaspari-Cohn closed form (value 1 at 0,
5/24 at r = c/2 from both pieces, 0 beyond c) and its integral against the
Riemann normalisation; the counts of frequencies, directions, centres and
coefficients from the formulas; a single isolated wavelet's field at its
centre (sqrt(2) for cos, 0 for sin); linearity and the time-factor sum.
"""
from __future__ import annotations

import numpy as np
from scipy.sparse import csc_matrix

KM2DEG = 1.0 / 110


def mywindow(x):
    """basis.py:4138-4140."""
    return np.cos(x * 0.5 * np.pi) ** 2


def gaspari_cohn(r, c=1.0):
    """tools.py:16-47 for an array or scalar r (scalars give a scalar)."""
    scalar = np.isscalar(r)
    ra = np.atleast_1d(np.asarray(r, dtype=np.float64))
    if c <= 0:
        out = np.zeros_like(ra)
    else:
        ra = 2 * np.abs(ra) / c
        out = np.zeros_like(ra)
        i = ra <= 1.0
        x = ra[i]
        out[i] = -0.25 * x ** 5 + 0.5 * x ** 4 + 0.625 * x ** 3 - 5.0 / 3.0 * x ** 2 + 1.0
        i = (ra > 1.0) & (ra <= 2.0)
        x = ra[i]
        out[i] = 1.0 / 12.0 * x ** 5 - 0.5 * x ** 4 + 0.625 * x ** 3 + 5.0 / 3.0 * x ** 2 - 5.0 * x + 4.0 \
            - 2.0 / 3.0 / x
    return float(out[0]) if scalar else out


def _ladder(lo, hi, step):
    """Ascending pseudo-frequencies 1/lo .. down to 1/hi (basis.py:2698-2714)."""
    logf = np.arange(np.log(1.0 / lo), np.log(1.0 / hi) - step, -step)[::-1]
    f = np.exp(logf)
    return f[1.0 / f <= hi]


def _centred(mid, lo, hi, spacing):
    """Values spaced ``spacing`` around ``mid``, down to > lo and up to < hi
    (the arange pairs of basis.py:2727-2747, 2774-2783)."""
    left = np.arange(mid, lo, -spacing)[::-1]
    right = np.arange(mid, hi, spacing)[1:]
    return np.concatenate((left, right))


class BasisWavelet3dOracle:
    """``Basis_wavelet3d`` on a grid given by 2-D ``lon``, ``lat`` (degrees)
    and an optional boolean ``mask`` (True = masked)."""

    controls_initial_state = False
    time_dependent = True

    def __init__(self, lon, lat, mask=None, facnst=1.0, npsp=3.5, facpsp=1.5, lmin=80.0, lmax=970.0,
                 tmin=2.0, tmax=20.0, sigma_Q=1e-1):
        self.lon1d = np.asarray(lon, dtype=np.float64).ravel()
        self.lat1d = np.asarray(lat, dtype=np.float64).ravel()
        self.shape_phys = np.shape(lon)
        self.nphys = self.lon1d.size
        self.lon_min, self.lon_max = self.lon1d.min(), self.lon1d.max()
        self.lat_min, self.lat_max = self.lat1d.min(), self.lat1d.max()
        m = None if mask is None else np.asarray(mask, dtype=bool).ravel()
        self.mask1d = m if (m is not None and np.any(m)) else None
        self.facnst, self.npsp, self.facpsp = facnst, npsp, facpsp
        self.lmin, self.lmax, self.tmin, self.tmax, self.sigma_Q = lmin, lmax, tmin, tmax, sigma_Q

    def set_basis(self, time, return_q=False):
        time = np.asarray(time, dtype=np.float64)
        self.time = time
        tmin_, tmax_ = time.min(), time.max()
        lon_min, lon_max = self.lon_min, self.lon_max
        if lon_max < lon_min:
            lon_max = lon_max + 360.0
        lat_min, lat_max = self.lat_min, self.lat_max
        step = np.log(1 + self.facpsp / self.npsp)
        fs = _ladder(self.lmin, self.lmax, step)
        dfs = fs[1:] - fs[:-1]
        nfs = len(fs)
        ft = _ladder(self.tmin, self.tmax, step)
        nft = len(ft)
        theta = np.linspace(0, np.pi, int(np.pi * fs[0] / dfs[0] * self.facpsp))[:-1]
        ntheta = len(theta)
        DXs = 1.0 / fs * self.npsp * 0.5
        NS = np.empty(nfs, dtype=np.int64)
        ENSLON, ENSLAT = [None] * nfs, [None] * nfs
        for ifs in range(nfs):
            lats = _centred((lat_min + lat_max) / 2, lat_min - DXs[ifs] * KM2DEG, lat_max + DXs[ifs] * KM2DEG,
                            DXs[ifs] / self.facnst * KM2DEG)
            elon, elat = [], []
            for la in lats:
                cl = np.cos(la * np.pi / 180.0)
                lons = _centred((lon_min + lon_max) / 2, lon_min - DXs[ifs] / cl * KM2DEG,
                                lon_max + DXs[ifs] / cl * KM2DEG, DXs[ifs] / self.facnst / cl * KM2DEG)
                for lo in lons:
                    if self.mask1d is not None:
                        ind = np.where((np.abs((self.lon1d - lo) / KM2DEG * cl) <= 0.5 / fs[ifs])
                                       & (np.abs((self.lat1d - la) / KM2DEG) <= 0.5 / fs[ifs]))[0]
                        if np.all(self.mask1d[ind]):
                            continue
                    elon.append(lo)
                    elat.append(la)
            ENSLON[ifs] = np.asarray(elon, dtype=np.float64)
            ENSLAT[ifs] = np.asarray(elat, dtype=np.float64)
            NS[ifs] = len(elon)
        NT = np.empty(nft, dtype=np.int64)
        DXt = 1.0 / ft * self.npsp * 0.5
        ENST, norm_fact = [None] * nft, [None] * nft
        for ift in range(nft):
            ENST[ift] = _centred((tmin_ + tmax_) / 2, tmin_ - DXt[ift], tmax_ + DXt[ift], DXt[ift] / self.facnst)
            NT[ift] = ENST[ift].size
            tt = np.linspace(-DXt[ift], DXt[ift])
            acc = np.zeros_like(tt)
            for i in range(tt.size - 1):
                acc[i + 1] = acc[i] + gaspari_cohn(tt[i], DXt[ift]) * (tt[i + 1] - tt[i])
            norm_fact[ift] = acc.max()
        bounds = [0] * (nfs + 1)
        iwave = 0
        for iff in range(nfs):
            bounds[iff] = iwave
            if NS[iff] > 0:
                iwave += 2 * int(NT.sum()) * ntheta * int(NS[iff])
        bounds[-1] = iwave
        self.fs, self.ft, self.theta, self.ntheta, self.nfs, self.nft = fs, ft, theta, ntheta, nfs, nft
        self.DXs, self.DXt, self.NS, self.NT = DXs, DXt, NS, NT
        self.ENSLON, self.ENSLAT, self.ENST, self.norm_fact = ENSLON, ENSLAT, ENST, norm_fact
        self.fs_wavebounds = bounds
        self.nbasis = iwave
        self.k = 2 * np.pi * fs
        self.Gx, self.Nx = self._space()
        self.Gt, self.Nt = self._time(time)
        if return_q:
            return np.zeros(self.nbasis), self.sigma_Q * np.ones(self.nbasis)

    def _space(self):
        Gx, Nx = [None] * self.nfs, [None] * self.nfs
        for iff in range(self.nfs):
            data, rows, sizes = [], [], []
            for P in range(self.NS[iff]):
                lo, la = self.ENSLON[iff][P], self.ENSLAT[iff][P]
                cl = np.cos(la * np.pi / 180.0)
                ind = np.where((np.abs((self.lon1d - lo) / KM2DEG * cl) <= self.DXs[iff])
                               & (np.abs((self.lat1d - la) / KM2DEG) <= self.DXs[iff]))[0]
                xx = (self.lon1d[ind] - lo) / KM2DEG * cl
                yy = (self.lat1d[ind] - la) / KM2DEG
                if self.mask1d is not None:
                    keep = ~self.mask1d[ind]
                    ind, xx, yy = ind[keep], xx[keep], yy[keep]
                facs = mywindow(xx / self.DXs[iff]) * mywindow(yy / self.DXs[iff])
                for th in self.theta:
                    kx = self.k[iff] * np.cos(th)
                    ky = self.k[iff] * np.sin(th)
                    ph = kx * xx + ky * yy
                    for val in (np.sqrt(2) * facs * np.cos(ph), np.sqrt(2) * facs * np.sin(ph)):
                        rows.append(ind)
                        data.append(val)
                        sizes.append(ind.size)
            nw = len(sizes)
            Nx[iff] = nw
            indptr = np.zeros(nw + 1, dtype=np.int64)
            indptr[1:] = np.cumsum(sizes)
            Gx[iff] = csc_matrix((np.concatenate(data) if data else np.zeros(0),
                                  np.concatenate(rows) if rows else np.zeros(0, dtype=np.int64), indptr),
                                 shape=(self.nphys, nw))
        return Gx, Nx

    def _time(self, time):
        Gt, Nt = [], []
        for t in time:
            gt, nt = [None] * self.nfs, [0] * self.nfs
            for ifs in range(self.nfs):
                g = np.zeros(self.fs_wavebounds[ifs + 1] - self.fs_wavebounds[ifs])
                blk = 2 * self.ntheta * int(self.NS[ifs])
                pos = 0
                for ift in range(self.nft):
                    for tc in self.ENST[ift]:
                        dt = t - tc
                        if abs(dt) < self.DXt[ift]:
                            fact = gaspari_cohn(dt, self.DXt[ift]) / self.norm_fact[ift]
                            if fact != 0:
                                nt[ifs] += 1
                                g[pos:pos + blk] = fact
                        pos += blk
                gt[ifs] = g
            Gt.append(gt)
            Nt.append(nt)
        return Gt, Nt

    def operg(self, it, X):
        """Field at basis time index ``it`` (basis.py:2936-2960)."""
        X = np.asarray(X)
        phi = np.zeros(self.nphys, dtype=np.result_type(X.dtype, np.float64))
        for iff in range(self.nfs):
            b0, b1 = self.fs_wavebounds[iff], self.fs_wavebounds[iff + 1]
            g = self.Gt[it][iff]
            ind0 = np.nonzero(g)[0]
            if ind0.size > 0:
                gx = (g * X[b0:b1])[ind0].reshape(self.Nt[it][iff], self.Nx[iff])
                phi += self.Gx[iff].dot(gx.sum(axis=0))
        return phi.reshape(self.shape_phys)

    def operg_transpose(self, it, adphi):
        """Coefficient gradient of a field cotangent at basis time index ``it``
        (basis.py:2962-2985)."""
        adX = np.zeros(self.nbasis)
        ad = np.asarray(adphi, dtype=np.float64).ravel()
        for iff in range(self.nfs):
            b0, b1 = self.fs_wavebounds[iff], self.fs_wavebounds[iff + 1]
            g = self.Gt[it][iff]
            ind0 = np.nonzero(g)[0]
            if ind0.size > 0:
                gg = g[ind0].reshape(self.Nt[it][iff], self.Nx[iff])
                a = self.Gx[iff].T.dot(ad)
                a = np.repeat(a[np.newaxis, :], self.Nt[it][iff], axis=0)
                adX[b0:b1][ind0] += (gg * a).ravel()
        return adX
