"""Seeded synthetic inputs for closed-basin QG 4D-Var problems.

Builds what the reference reads from files for a real experiment: the grid
(``State`` with lon/lat, DX, DY; mapping/src/state.py:137-250), an initial SSH
field, and along-track / swath observations (``dict_obs``;
mapping/src/obs.py:20-330) sampled from a truth run.

* Grid: constant spacing in km around a centre point, laid out like
  ``State.ini_car_grid_center`` (state.py:221-260).
* Initial SSH: isotropic Gaussian random field with an energy spectrum
  ``k exp(-(k - k0)^2 / (2 (k0/2)^2))`` peaked at ``2 pi / wavelength``,
  scaled to the requested RMS and tapered to zero over the outer 8 cells so
  that the field is compatible with psi = 0 on the boundary.
* Nadir tracks: straight lines at seeded angles and offsets, sampled every
  ``spacing_km``, each observed at one seeded model step.
* Swaths: one straight band per day, ``width_km`` wide, pixels every
  ``pixel_km`` along and across track, at a seeded step of that day.
* Observation values: the truth SSH at the observation step, interpolated
  bilinearly in grid-index space, plus seeded N(0, sigma) noise.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

KM_PER_DEG = 111.32


@dataclass
class Grid:
    ny: int
    nx: int
    dx: float  # metres
    dy: float  # metres
    lon: np.ndarray
    lat: np.ndarray
    lon_center: float
    lat_center: float

    @property
    def DX(self):
        return np.full((self.ny, self.nx), self.dx)

    @property
    def DY(self):
        return np.full((self.ny, self.nx), self.dy)

    def index_to_lonlat(self, iy, ix):
        """Map fractional grid indices to lon/lat with the grid's own layout."""
        lat = self.lat_center + (np.asarray(iy) - self.ny // 2) * (self.dy * 1e-3 / KM_PER_DEG)
        lon = self.lon_center + (np.asarray(ix) - self.nx // 2) * (
            self.dx * 1e-3 / (KM_PER_DEG * np.cos(np.radians(lat))))
        return lon, lat


def make_grid(ny, nx, dx_m, dy_m=None, lon_center=-60.0, lat_center=35.0) -> Grid:
    dy_m = dx_m if dy_m is None else dy_m
    iy, ix = np.meshgrid(np.arange(ny, dtype=float), np.arange(nx, dtype=float), indexing="ij")
    g = Grid(ny, nx, float(dx_m), float(dy_m), None, None, lon_center, lat_center)
    g.lon, g.lat = g.index_to_lonlat(iy, ix)
    return g


def gaussian_random_field(ny, nx, dx_m, wavelength_m, rms, seed, taper=8, dy_m=None):
    """Isotropic Gaussian random field with a spectral peak at ``wavelength_m``."""
    dy_m = dx_m if dy_m is None else dy_m
    rng = np.random.default_rng(seed)
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=dy_m)
    kx = 2 * np.pi * np.fft.fftfreq(nx, d=dx_m)
    KY, KX = np.meshgrid(ky, kx, indexing="ij")
    k = np.sqrt(KX ** 2 + KY ** 2)
    k0 = 2 * np.pi / wavelength_m
    spec = k * np.exp(-(k - k0) ** 2 / (2 * (0.5 * k0) ** 2))
    amp = np.sqrt(spec / np.maximum(k, 1e-30))
    noise = rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))
    fld = np.real(np.fft.ifft2(amp * noise))
    wy = np.ones(ny)
    wx = np.ones(nx)
    r = np.arange(taper)
    ramp = 0.5 - 0.5 * np.cos(np.pi * np.clip(r - 2, 0, None) / max(taper - 2, 1))
    ramp[:3] = 0.0
    wy[:taper] = ramp
    wy[-taper:] = ramp[::-1]
    wx[:taper] = ramp
    wx[-taper:] = ramp[::-1]
    fld = fld * wy[:, None] * wx[None, :]
    fld *= rms / np.sqrt(np.mean(fld ** 2))
    return fld


@dataclass
class ObsSet:
    """Observations grouped by model step (the reference's ``dict_obs``)."""

    steps: dict = field(default_factory=dict)  # step -> dict(iy, ix, lon, lat, y, err)

    def add(self, step, iy, ix, lon, lat, err):
        e = self.steps.setdefault(step, dict(iy=[], ix=[], lon=[], lat=[], err=[]))
        e["iy"].append(iy)
        e["ix"].append(ix)
        e["lon"].append(lon)
        e["lat"].append(lat)
        e["err"].append(err)

    def finalize(self):
        for e in self.steps.values():
            for k in ("iy", "ix", "lon", "lat", "err"):
                if isinstance(e[k], list):
                    e[k] = np.concatenate(e[k]) if e[k] else np.zeros(0)
        self.steps = dict(sorted(self.steps.items()))
        return self

    @property
    def n_obs(self):
        return int(sum(e["iy"].size for e in self.steps.values()))


def _line_points(grid: Grid, angle, offset_frac, spacing_m, margin=2.0):
    """Points of a straight line through the basin, ``spacing_m`` apart, as
    fractional (iy, ix) inside ``[margin, n-1-margin]``."""
    Ly = (grid.ny - 1) * grid.dy
    Lx = (grid.nx - 1) * grid.dx
    cy, cx = 0.5 * Ly, 0.5 * Lx
    d = np.array([np.sin(angle), np.cos(angle)])
    nrm = np.array([d[1], -d[0]])
    half = 0.5 * np.hypot(Lx, Ly)
    p0 = np.array([cy, cx]) + (offset_frac - 0.5) * 2 * half * nrm
    s = np.arange(-half, half + spacing_m, spacing_m)
    y = p0[0] + s * d[0]
    x = p0[1] + s * d[1]
    iy, ix = y / grid.dy, x / grid.dx
    keep = (iy >= margin) & (iy <= grid.ny - 1 - margin) & (ix >= margin) & (ix <= grid.nx - 1 - margin)
    return iy[keep], ix[keep], d, nrm


def nadir_tracks(grid: Grid, n_tracks, spacing_km, n_steps, sigma, seed, obs: ObsSet | None = None,
                 step_seed=None):
    """``n_tracks`` straight tracks at seeded angles/offsets (``seed``) observed at
    seeded steps in [1, n_steps] (``step_seed``, default ``seed``)."""
    rng = np.random.default_rng(seed)
    rng_step = np.random.default_rng(seed if step_seed is None else step_seed)
    obs = ObsSet() if obs is None else obs
    for _ in range(n_tracks):
        angle = rng.uniform(0, np.pi)
        off = rng.uniform(0.2, 0.8)
        step = int(rng_step.integers(1, n_steps + 1))
        iy, ix, _, _ = _line_points(grid, angle, off, spacing_km * 1e3)
        if iy.size == 0:
            continue
        lon, lat = grid.index_to_lonlat(iy, ix)
        obs.add(step, iy, ix, lon, lat, np.full(iy.size, sigma))
    return obs


def swaths(grid: Grid, n_days, steps_per_day, width_km, pixel_km, sigma, seed, obs: ObsSet | None = None,
           step_seed=None, max_step=None):
    """One straight swath per day, ``width_km`` wide with ``pixel_km`` pixels;
    swaths after ``max_step`` (a window shorter than a day) are dropped."""
    rng = np.random.default_rng(seed)
    rng_step = np.random.default_rng(seed if step_seed is None else step_seed)
    obs = ObsSet() if obs is None else obs
    n_across = int(round(width_km / pixel_km))
    across = (np.arange(n_across) - 0.5 * (n_across - 1)) * pixel_km * 1e3
    for day in range(n_days):
        angle = rng.uniform(0, np.pi)
        off = rng.uniform(0.3, 0.7)
        step = int(day * steps_per_day + rng_step.integers(1, steps_per_day + 1))
        if max_step is not None and step > max_step:
            continue
        iy0, ix0, d, nrm = _line_points(grid, angle, off, pixel_km * 1e3, margin=-1e9)
        y = iy0[:, None] * grid.dy + across[None, :] * nrm[0]
        x = ix0[:, None] * grid.dx + across[None, :] * nrm[1]
        iy, ix = (y / grid.dy).ravel(), (x / grid.dx).ravel()
        keep = (iy >= 2) & (iy <= grid.ny - 3) & (ix >= 2) & (ix <= grid.nx - 3)
        iy, ix = iy[keep], ix[keep]
        if iy.size == 0:
            continue
        lon, lat = grid.index_to_lonlat(iy, ix)
        obs.add(step, iy, ix, lon, lat, np.full(iy.size, sigma))
    return obs


def bilinear(field2d, iy, ix):
    """Bilinear interpolation of ``field2d`` at fractional grid indices."""
    ny, nx = field2d.shape
    i0 = np.clip(np.floor(iy).astype(np.int64), 0, ny - 2)
    j0 = np.clip(np.floor(ix).astype(np.int64), 0, nx - 2)
    wy = iy - i0
    wx = ix - j0
    return ((1 - wy) * (1 - wx) * field2d[i0, j0] + (1 - wy) * wx * field2d[i0, j0 + 1] +
            wy * (1 - wx) * field2d[i0 + 1, j0] + wy * wx * field2d[i0 + 1, j0 + 1])


def sample_truth(obs: ObsSet, truth_at_step, sigma, seed):
    """Fill ``y`` with the truth at each observation step plus seeded noise.

    ``truth_at_step(steps)`` returns a dict step -> (ny, nx) SSH field.
    """
    rng = np.random.default_rng(seed)
    fields = truth_at_step(sorted(obs.steps))
    for s, e in obs.steps.items():
        e["y"] = bilinear(np.asarray(fields[s], dtype=np.float64), e["iy"], e["ix"]) + \
            sigma * rng.standard_normal(e["iy"].size)
    return obs


# ----------------------------------------------------------------------------
# Named workloads
# ----------------------------------------------------------------------------
@dataclass
class QgCase:
    name: str
    ny: int
    nx: int
    dx: float
    c: float
    dt: float
    n_steps: int
    f0: float = 1e-4
    g: float = 9.81
    time_scheme: str = "rk2"
    rms: float = 0.1
    wavelength: float = 100e3
    n_tracks: int = 10
    track_km: float = 7.0
    swath_per_day: bool = False
    swath_width_km: float = 120.0
    swath_pixel_km: float = 2.0
    sigma_obs: float = 0.02
    sigma_b: float = 0.1
    checkpoint_steps: int = 36
    seed: int = 1234

    @property
    def days(self):
        return self.n_steps * self.dt / 86400.0


def case_basin128():
    """128x128, dx=10 km, c=2.7 m/s, dt=1200 s, 7 days (504 steps), 10 tracks @7 km."""
    return QgCase("qg128_7d", 128, 128, 10e3, 2.7, 1200.0, 504, n_tracks=10, track_km=7.0,
                  checkpoint_steps=36)


def case_basin256():
    """256x256, dx=10 km, c=2.7 m/s, dt=1200 s, 14 days (1008 steps)."""
    return QgCase("qg256_14d", 256, 256, 10e3, 2.7, 1200.0, 1008, n_tracks=20, track_km=7.0,
                  checkpoint_steps=32)


def case_basin512():
    """512x512, dx=5 km, Ld=20 km (c=2 m/s), dt=900 s, 30 days (2880 steps),
    60 tracks @6 km + one 120 km swath per day at 2 km pixels, sigma=0.02 m,
    checkpoints every 54 steps (about sqrt(2880))."""
    return QgCase("qg512_30d", 512, 512, 5e3, 2.0, 900.0, 2880, n_tracks=60, track_km=6.0,
                  swath_per_day=True, checkpoint_steps=54)


def make_observations(case: QgCase, grid: Grid, n_steps=None, member=0):
    """Tracks (and swaths) of ``case``.  Batch members share the observation
    steps (the pass times) and differ in track geometry."""
    n_steps = case.n_steps if n_steps is None else n_steps
    gseed = case.seed + 1 + 1009 * member
    obs = nadir_tracks(grid, case.n_tracks, case.track_km, n_steps, case.sigma_obs, gseed,
                       step_seed=case.seed + 1)
    if case.swath_per_day:
        spd = int(round(86400 / case.dt))
        n_days = max(1, int(n_steps // spd))
        swaths(grid, n_days, spd, case.swath_width_km, case.swath_pixel_km, case.sigma_obs,
               case.seed + 2 + 1009 * member, obs, step_seed=case.seed + 2, max_step=n_steps)
    return obs.finalize()


def config4_cases(n=48, seed=2024, dx=5e3, dt=900.0):
    """``n`` independent regional problems of uneven size and window length:
    sides drawn from {96, 128, 192, 256, 320, 384} (squares and rectangles),
    windows uniform in 5-20 days at dt = 900 s, Ld uniform in 15-40 km
    (c = f0 Ld), initial-field RMS uniform in 0.05-0.15 m, 5-40 nadir tracks
    every 7 km; checkpoints about every sqrt(window) steps.  Each problem has
    its own seed."""
    rng = np.random.default_rng(seed)
    sides = np.array([96, 128, 192, 256, 320, 384])
    f0 = 1e-4
    cases = []
    for p in range(n):
        ny, nx = (int(v) for v in rng.choice(sides, 2))
        days = rng.uniform(5.0, 20.0)
        n_steps = int(round(days * 86400.0 / dt))
        ld = rng.uniform(15e3, 40e3)
        cases.append(QgCase(f"c4_{p:02d}_{ny}x{nx}_{n_steps}", ny, nx, dx, f0 * ld, dt, n_steps,
                            f0=f0, rms=float(rng.uniform(0.05, 0.15)),
                            n_tracks=int(rng.integers(5, 41)), track_km=7.0,
                            checkpoint_steps=int(round(np.sqrt(n_steps))),
                            seed=int(seed + 7919 * (p + 1))))
    return cases
