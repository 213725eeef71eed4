"""Control-to-model bases with the reference's ``Basis`` protocol
(``set_basis``, ``operg``, ``operg_transpose``, ``nbasis``;
mapping/src/basis.py:30-86, 3939-4033, datlas-ocean/VarDyn).

* ``Basis_offset``  -- spatially constant SSH flux (basis.py:3939-4000).
* ``Basis_ic``      -- the initial SSH itself (no reference class; the
  reference's reverse sweep computes this adjoint and drops it).

``operg`` / ``operg_transpose`` act on device tensors in ``State.params`` /
``State.var`` dictionaries.
"""
from __future__ import annotations

import numpy as np
import torch


class Basis_offset:
    controls_initial_state = False

    def __init__(self, State, sigma_B, name_mod_var="ssh"):
        self.name_mod_var = name_mod_var
        self.shape_phys = (State.ny, State.nx)
        self.sigma_B = sigma_B
        self.nbasis = 1

    def set_basis(self, time=None, return_q=False, **kwargs):
        Q = self.sigma_B * np.ones(self.nbasis)
        if return_q:
            return np.zeros_like(Q), Q

    def operg(self, t, X, State=None):
        Xt = torch.as_tensor(X, dtype=torch.float64).reshape(-1)
        if State is None:
            return float(Xt[0]) * np.ones(self.shape_phys)
        ref = State[self.name_mod_var]
        State[self.name_mod_var] = Xt[0].to(ref.device, ref.dtype).expand_as(ref).contiguous().clone()

    def operg_transpose(self, t, adState):
        ad = adState[self.name_mod_var]
        out = ad.double().sum().reshape(1)
        adState[self.name_mod_var] = torch.zeros_like(ad)
        return out.cpu().numpy()


class Basis_ic:
    controls_initial_state = True

    def __init__(self, State, sigma_B, name_mod_var="ssh", land=None):
        self.name_mod_var = name_mod_var
        self.shape_phys = (State.ny, State.nx)
        self.sigma_B = sigma_B
        self.nbasis = State.ny * State.nx
        self.land = None if land is None else torch.as_tensor(land, device=State.device)

    def set_basis(self, time=None, return_q=False, **kwargs):
        Q = self.sigma_B * np.ones(self.nbasis)
        if return_q:
            return np.zeros_like(Q), Q

    def operg(self, t, X, State=None):
        v = torch.as_tensor(np.asarray(X, dtype=np.float64)).reshape(self.shape_phys)
        if State is None:
            return v.numpy()
        ref = State[self.name_mod_var]
        v = v.to(ref.device, ref.dtype).contiguous()
        if self.land is not None:
            v[self.land] = float("nan")
        State[self.name_mod_var] = v

    def operg_transpose(self, t, adState):
        return torch.nan_to_num(adState[self.name_mod_var].double()).reshape(-1).cpu().numpy()
