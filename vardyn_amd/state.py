"""Model state container with the ``var`` / ``params`` dictionaries of the
reference's ``State`` (mapping/src/state.py:25-656, datlas-ocean/VarDyn),
holding device tensors.

Only the protocol the 4D-Var path uses is kept: ``copy(free=...)``,
``getvar`` / ``setvar``, ``random``, ``scalar``, ``Sum`` (state.py:502-656).
Grid construction and netCDF I/O are not part of this package.
"""
from __future__ import annotations

import numpy as np
import torch


class State:
    def __init__(self, ny, nx, dtype=torch.float64, device=None, lon=None, lat=None, DX=None,
                 DY=None, mask=None, f=None, g=9.81):
        self.ny, self.nx = int(ny), int(nx)
        self.dtype = dtype
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None \
            else torch.device(device)
        self.lon, self.lat, self.DX, self.DY = lon, lat, DX, DY
        self.mask = mask
        self.f = f
        self.g = g
        self.var = {}
        self.params = {}

    def _new(self):
        s = State(self.ny, self.nx, self.dtype, self.device, self.lon, self.lat, self.DX, self.DY,
                  self.mask, self.f, self.g)
        return s

    def copy(self, free=False):
        """Deep copy; ``free=True`` gives zeros of the same layout (state.py:545-557)."""
        s = self._new()
        for d_src, d_dst in ((self.var, s.var), (self.params, s.params)):
            for k, v in d_src.items():
                d_dst[k] = torch.zeros_like(v) if free else v.clone()
        return s

    def getvar(self, name_var=None, vect=False):
        """Host copy of one variable (state.py:559-583), or all of them flattened."""
        if name_var is None:
            arrs = [v.double().cpu().numpy().ravel() for v in self.var.values()]
            return np.concatenate(arrs) if vect else arrs
        a = self.var[name_var].double().cpu().numpy()
        return a.ravel() if vect else a

    def getparams(self, name_params=None, vect=False):
        if name_params is None:
            arrs = [v.double().cpu().numpy().ravel() for v in self.params.values()]
            return np.concatenate(arrs) if vect else arrs
        a = self.params[name_params].double().cpu().numpy()
        return a.ravel() if vect else a

    def setvar(self, var, name_var):
        t = torch.as_tensor(np.asarray(var) if not isinstance(var, torch.Tensor) else var,
                            device=self.device).to(self.dtype)
        self.var[name_var] = t.reshape(self.ny, self.nx).contiguous().clone()

    def random(self, ampl=1):
        """State with uniform random var/params (state.py:609-628, seed 0)."""
        s = self.copy(free=True)
        np.random.seed(0)
        for d in (s.var, s.params):
            for k in d:
                d[k] = torch.as_tensor(ampl * np.random.random((self.ny, self.nx)),
                                       device=self.device).to(self.dtype)
        return s

    def scalar(self, coeff):
        for d in (self.var, self.params):
            for k in d:
                d[k] = d[k] * coeff

    def Sum(self, other):
        for d, o in ((self.var, other.var), (self.params, other.params)):
            for k in d:
                d[k] = d[k] + o[k]
