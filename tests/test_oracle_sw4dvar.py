"""Pins of the shallow-water 4D-Var cost oracle (oracle/sw4dvar.py).

* At rest without forcing the model stays at rest, so J is the closed form
  1/2 sum((0 - y)/err)^2 + 1/2 X.X/sigma^2 of the observations and the
  background term alone.
* From rest, one RK3 step with a forcing F gives dt F on the ocean points
  (the forcing pin of the SW oracle), so with observations one step after the
  start the cost is the closed form of the observed field dt phi/86400, phi
  the basis field of the first basis time: this fixes the day-to-second
  scale, the top-layer placement of the SSH forcing and the SSH as the sum of
  the layer thicknesses.
* The complex-step derivative of the cost equals centred finite differences.
"""
import numpy as np
import pytest

from oracle.basis_wavelet3d import KM2DEG, BasisWavelet3dOracle
from oracle.fourdvar import checkpoint_schedule
from oracle.obsop import ObsopOracle
from oracle.sw import SwOracle
from oracle.sw4dvar import SwCostOracle


def setup(n_steps=6, obs_steps=(1, 4), seed=0, rest=True):
    nx, ny, nl = 14, 12, 3
    dx, dt = 8e3, 120.0
    y = (np.arange(ny + 1) - ny / 2) * dx
    f = np.broadcast_to(1e-4 + 1.6e-11 * y, (nx + 1, ny + 1)).copy()
    sw = SwOracle(nx, ny, dx, dx, np.array([500.0, 1000.0, 3000.0]), np.array([0.02, 0.02, 0.01]), f, dt)
    lat1 = 43.0 + (np.arange(ny) - (ny - 1) / 2) * dx * 1e-3 * KM2DEG
    lon1 = 3.0 + (np.arange(nx) - (nx - 1) / 2) * dx * 1e-3 * KM2DEG / np.cos(np.radians(43.0))
    lat, lon = np.meshgrid(lat1, lon1)  # (nx, ny): the model's h layout
    rng = np.random.default_rng(seed)
    T = np.arange(n_steps + 1) * dt
    obs = {}
    for s in obs_steps:
        k = 9
        lo = rng.uniform(lon1[2], lon1[-3], k)
        la = rng.uniform(lat1[2], lat1[-3], k)
        obs[float(T[s])] = (lo, la, 0.05 * rng.standard_normal(k), np.full(k, 0.02))
    H = ObsopOracle(lon, lat, np.full(lon.shape, dx), np.full(lon.shape, dx), T, dt, obs, Npix=4)
    every = 3
    cps = checkpoint_schedule(n_steps + 1, every, lambda i: H.is_obs(T[i]))
    basis = BasisWavelet3dOracle(lon, lat, lmin=40.0, lmax=170.0, npsp=3.5, facpsp=3.5, tmin=0.05, tmax=0.05)
    basis.set_basis(np.arange(0, T[-1] + every * dt, every * dt) / 86400)
    if rest:
        x0 = (np.zeros((nl, nx + 1, ny)), np.zeros((nl, nx, ny + 1)), np.zeros((nl, nx, ny)))
    else:
        m = sw.m
        x0 = (0.05 * rng.standard_normal((nl, nx + 1, ny)) * m["u"],
              0.05 * rng.standard_normal((nl, nx, ny + 1)) * m["v"], 0.2 * rng.standard_normal((nl, nx, ny)) * m["h"])
    cost = SwCostOracle(sw, H, basis, cps, T, every, x0, sigma_b=0.1)
    return cost, H, basis, T, rng


def test_rest_without_forcing_closed_form():
    cost, H, basis, T, rng = setup()
    X0 = np.zeros(basis.nbasis)
    expect = 0.5 * sum(np.sum((H.y_arr[i, :H.n_obs[i]] / H.err_arr[i, :H.n_obs[i]]) ** 2)
                       for i in range(len(H.t_obs)))
    assert cost.cost(X0) == pytest.approx(expect, rel=1e-14)
    # the background term on its own: 1/2 X.X / sigma^2 for a control whose
    # forcing is never observed (observations only at the first step... of a
    # zero window there are none, so vary X and subtract)
    X1 = 0.01 * rng.standard_normal(basis.nbasis)
    c1 = SwCostOracle(cost.sw, H, basis, cost.cps, cost.T, cost.dtbasis, cost.x0, sigma_b=0.1)
    jb = 0.5 * np.dot(X1, X1) / 0.1 ** 2
    assert c1.cost(X1) - jb == pytest.approx(cost_without_background(c1, X1), rel=1e-12)


def cost_without_background(c, X):
    big = SwCostOracle(c.sw, c.H, c.basis, c.cps, c.T, c.dtbasis, c.x0, sigma_b=1e30)
    return big.cost(X)


def test_one_step_forced_observation_closed_form():
    cost, H, basis, T, rng = setup(obs_steps=(1,))
    X = 0.05 * rng.standard_normal(basis.nbasis)
    phi = basis.operg(0, X)                       # the basis field of the first basis time (per day)
    ssh1 = cost.sw.dt * phi / 86400.0             # one step from rest: dt F on the top layer
    m = (np.einsum("j,j->j", H.data_arr[0], ssh1.ravel()[H.col_arr[0]]))
    hx = np.zeros(H.y_arr.shape[1])
    np.add.at(hx, H.row_arr[0], m)
    mis = (hx - H.y_arr[0]) / H.err_arr[0]
    expect = 0.5 * (mis.dot(mis) + np.dot(X, X) / 0.1 ** 2)
    assert cost.cost(X) == pytest.approx(expect, rel=1e-12)


def test_complex_step_equals_finite_differences():
    cost, H, basis, T, rng = setup(rest=False)
    X = 0.05 * rng.standard_normal(basis.nbasis)
    d = rng.standard_normal(basis.nbasis)
    cs = cost.directional_derivative(X, d)
    eps = 1e-4
    fd = (cost.cost(X + eps * d) - cost.cost(X - eps * d)) / (2 * eps)
    assert cs == pytest.approx(fd, rel=1e-6)
