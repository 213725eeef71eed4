"""Assemble a synthetic 4D-Var problem of the multilayer shallow-water model
with a 3-D wavelet forcing control, on the GPU: what ``Inv_4Dvar`` sets up
before minimising (mapping/src/inv.py:105-263) for ``Model_qgsw`` with a
``Basis_wavelet3d`` control.

The workload of the 3-layer configuration: 256 x 256 cells of 8 km on a beta
plane (f = 1e-4 + 1.6e-11 y), layers 500 / 1000 / 3000 m, reduced gravities
0.02 (surface) / 0.02 / 0.01 m/s^2, dt = 120 s, a 10-day window; the control
is the SSH forcing of the top layer in a wavelet basis of 4 scales
(wavelengths 50-400 km, 9 directions) with 5-day time windows (55,404
coefficients on this grid); 30 seeded nadir tracks with sigma = 0.02 m.

Twin experiment: the initial state is a known balanced eddy field (the same
for truth and analysis); the truth is forced by a seeded coefficient vector
(scaled so that the forcing field has an RMS of ``forcing_rms`` m/day); the
observations sample the truth's SSH (the sum of the layer thicknesses) on
the tracks with seeded noise.
"""
from __future__ import annotations

from dataclasses import dataclass
from types import SimpleNamespace

import numpy as np
import torch

from . import synthetic as syn
from .basis_wavelet3d import Basis_wavelet3d
from .inv import checkpoint_schedule
from .obsop import Obsop_interp_l3
from .sw import SW
from .sw4dvar import SwFourDVar


@dataclass
class SwCase:
    name: str = "sw3l_256"
    nx: int = 256
    ny: int = 256
    dx: float = 8e3
    f0: float = 1e-4
    beta: float = 1.6e-11
    H: tuple = (500.0, 1000.0, 3000.0)
    g_prime: tuple = (0.02, 0.02, 0.01)
    dt: float = 120.0
    days: float = 10.0
    n_tracks: int = 30
    track_spacing_km: float = 7.0
    sigma_obs: float = 0.02
    lmin: float = 50.0
    lmax: float = 420.0
    npsp: float = 3.5
    facpsp: float = 3.5
    tmin: float = 5.0
    tmax: float = 5.0
    sigma_Q: float = 0.1
    checkpoint_steps: int = 720     # basis times and checkpoints every day
    eddy_rms: float = 0.05
    eddy_scale_m: float = 50e3
    forcing_rms: float = 0.05       # m/day
    seed: int = 2024
    visc_coef: float = 0.0          # Laplacian viscosity of u, v (m^2/s; the reference's default 0)
    diff_coef: float = 0.0          # Laplacian diffusion of h (m^2/s)

    @property
    def n_steps(self):
        return int(round(self.days * 86400 / self.dt))

    @property
    def lat0(self):
        return float(np.degrees(np.arcsin(self.f0 / (2 * 7.2921e-5))))


def small_case(**kw) -> SwCase:
    """A grid and window the NumPy oracle evaluates in seconds."""
    base = dict(name="sw3l_small", nx=40, ny=36, days=240 * 120.0 / 86400, n_tracks=6, lmin=40.0, lmax=170.0,
                tmin=0.1, tmax=0.1, checkpoint_steps=60, eddy_scale_m=30e3, forcing_rms=0.5)
    base.update(kw)
    return SwCase(**base)


def sw_params(case: SwCase):
    nx, ny = case.nx, case.ny
    y = (np.arange(ny + 1) - ny / 2) * case.dx
    f = np.broadcast_to(case.f0 + case.beta * y, (nx + 1, ny + 1)).copy()
    nl = len(case.H)
    return dict(nx=nx, ny=ny, nl=nl, dx=case.dx, dy=case.dx, H=np.asarray(case.H).reshape(nl, 1, 1),
                g_prime=np.asarray(case.g_prime), f=f, dt=case.dt, slip_coef=0.0, bottom_drag_coef=0.0,
                visc_coef=case.visc_coef, diff_coef=case.diff_coef)


def eddy_state(case: SwCase, param):
    """A seeded smooth thickness anomaly in the top layer (Gaussian smoothing of
    ``eddy_scale_m``, RMS ``eddy_rms``) with the geostrophic velocities of its
    pressure, zero in the layers below (SW layout)."""
    from scipy.ndimage import gaussian_filter
    nx, ny, nl = case.nx, case.ny, len(case.H)
    f, gp = param["f"], param["g_prime"]
    rng = np.random.default_rng(case.seed)
    eta = gaussian_filter(rng.standard_normal((nx, ny)), case.eddy_scale_m / case.dx)
    eta *= case.eddy_rms / eta.std()
    h = np.zeros((nl, nx, ny)); h[0] = eta
    f_u = 0.5 * (f[:, 1:] + f[:, :-1]); f_v = 0.5 * (f[1:, :] + f[:-1, :])
    u = np.zeros((nl, nx + 1, ny)); v = np.zeros((nl, nx, ny + 1))
    ey = np.gradient(eta, case.dx, axis=1); ex = np.gradient(eta, case.dx, axis=0)
    u[0, 1:nx, :] = -gp[0] * 0.5 * (ey[1:] + ey[:-1]) / f_u[1:nx]
    v[0, :, 1:ny] = gp[0] * 0.5 * (ex[:, 1:] + ex[:, :-1]) / f_v[:, 1:ny]
    return u, v, h


@dataclass
class SwProblem:
    case: SwCase
    grid: syn.Grid
    model: SW
    basis: Basis_wavelet3d
    H: Obsop_interp_l3
    T: np.ndarray
    checkpoints: np.ndarray
    time_basis: np.ndarray
    x0: tuple
    X_true: np.ndarray
    Q: np.ndarray
    dict_obs: dict
    obs: syn.ObsSet
    engine: SwFourDVar | None = None


def basis_kwargs(case: SwCase):
    return dict(lmin=case.lmin, lmax=case.lmax, npsp=case.npsp, facpsp=case.facpsp, tmin=case.tmin,
                tmax=case.tmax, sigma_Q=case.sigma_Q)


def _forced_run(model, basis, x0, X, steps, case: SwCase, n_steps):
    """SSH (SW layout) of the forced run at ``steps``: the basis field at day
    boundaries k * checkpoint_steps, held until the next one."""
    dev = model.device
    nb = n_steps // case.checkpoint_steps + 1
    times = np.arange(nb) * case.checkpoint_steps * case.dt / 86400
    fields = basis.operg_batch(times, X)
    nl = model.nl
    u, v, h = (torch.as_tensor(a, device=dev) for a in x0)
    out, s0 = {}, 0
    targets = sorted(set(int(s) for s in steps) | {k * case.checkpoint_steps for k in range(nb)
                                                   if k * case.checkpoint_steps <= n_steps})
    for s in targets:
        while s0 < s:
            k = s0 // case.checkpoint_steps
            n = min(s, (k + 1) * case.checkpoint_steps) - s0
            Fh = torch.zeros((nl, model.nx, model.ny), dtype=torch.float64, device=dev)
            Fh[0] = fields[k] / 86400.0
            u, v, h = model.step(u, v, h, nstep=n, Fh=Fh)
            s0 += n
        out[s] = h.sum(0).cpu().numpy()
    return out


def build_sw_problem(case: SwCase | None = None, device=None, with_engine=True, prec=False,
                     tape="auto") -> SwProblem:
    case = SwCase() if case is None else case
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    n_steps = case.n_steps
    param = sw_params(case)
    model = SW(param, device=dev)
    x0 = eddy_state(case, param)
    # the State grid (ny, nx); model, basis and observation tables use its transpose (nx, ny)
    grid = syn.make_grid(case.ny, case.nx, case.dx, lat_center=case.lat0)
    lon_sw, lat_sw = grid.lon.T.copy(), grid.lat.T.copy()
    T = np.arange(n_steps + 1) * case.dt
    step_basis = case.checkpoint_steps * case.dt
    # inv.py:184: the basis times, every checkpoint interval, in days
    time_basis = np.arange(0, T[-1] + step_basis, step_basis) / 24 / 3600
    basis = Basis_wavelet3d(SimpleNamespace(lon=lon_sw, lat=lat_sw, mask=None), device=dev, **basis_kwargs(case))
    _, Q = basis.set_basis(time_basis, return_q=True)
    # truth forcing: seeded coefficients scaled to the requested field RMS
    rng = np.random.default_rng(case.seed + 17)
    X_true = rng.standard_normal(basis.nbasis) * Q
    rms = float(torch.sqrt(torch.mean(basis.operg_batch(time_basis, X_true) ** 2)))
    X_true *= case.forcing_rms / rms
    obs = syn.nadir_tracks(grid, case.n_tracks, case.track_spacing_km, n_steps, case.sigma_obs, case.seed + 5)
    obs.finalize()
    fields = _forced_run(model, basis, x0, X_true, sorted(obs.steps), case, n_steps)
    syn.sample_truth(obs, lambda steps: {s: fields[s].T for s in steps}, case.sigma_obs, case.seed + 3)
    dict_obs = {float(T[s]): (e["lon"], e["lat"], e["y"], e["err"]) for s, e in obs.steps.items()}
    H = Obsop_interp_l3(lon_sw, lat_sw, np.full(lon_sw.shape, case.dx), np.full(lon_sw.shape, case.dx), T,
                        case.dt, dict_obs, Npix=4, device=dev).process_obs()
    cps = checkpoint_schedule(n_steps + 1, case.checkpoint_steps, lambda i: H.is_obs(T[i]))
    pb = SwProblem(case, grid, model, basis, H, T, cps, time_basis, x0, X_true, Q, dict_obs, obs)
    if with_engine:
        pb.engine = SwFourDVar(model, H, cps, T, basis, x0, sigma_b=Q, prec=prec, dtbasis=case.checkpoint_steps,
                               tape=tape)
    return pb
