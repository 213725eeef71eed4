"""Control-to-model bases with the reference's ``Basis`` protocol
(``set_basis``, ``operg``, ``operg_transpose``, ``nbasis``;
mapping/src/basis.py:30-86, 3939-4033, datlas-ocean/VarDyn).

* ``Basis``         -- the reference's factory (basis.py:30-86): the class
  named by ``config.BASIS.super``.
* ``Basis_offset``  -- spatially constant SSH flux (basis.py:3939-4000).
* ``Basis_ic``      -- the initial SSH itself (no reference class; the
  reference's reverse sweep computes this adjoint and drops it).

``operg`` / ``operg_transpose`` act on device tensors in ``State.params`` /
``State.var`` dictionaries.
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import VardynError


def Basis(config, State, verbose=True, multi_mode=False, **kwargs):
    """basis.py:30-86: the reduced basis named by ``config.BASIS.super``, or
    None without a basis.  The GPU classes stand in for both the NumPy and the
    JAX class of the reference (``BASIS_BM`` / ``BASIS_BM_JAX``,
    ``BASIS_OFFSET`` / ``BASIS_OFFSET_JAX``; the same operators); bases not
    built here (Gaussian, auxiliary-BM, boundary-condition, multi-basis) raise
    VardynError where the reference exits with 'not implemented'.
    ``multi_mode`` accumulation is not supported."""
    B = getattr(config, "BASIS", None)
    if B is None:
        return None
    sup = getattr(B, "super", None)
    if multi_mode:
        raise VardynError("Basis: multi_mode is not supported")
    if sup is None:
        raise VardynError("Basis: the multi-basis form (BASIS.super None) is not built")
    if verbose:
        print(B)
    if sup in ("BASIS_BM", "BASIS_BM_JAX"):
        from .basis_bm import Basis_bm
        return Basis_bm.from_config(config, State, **kwargs)
    if sup in ("BASIS_WAVELET3D", "BASIS_WAVELET3D_JAX"):
        from .basis_wavelet3d import Basis_wavelet3d
        return Basis_wavelet3d.from_config(config, State, **kwargs)
    if sup in ("BASIS_OFFSET", "BASIS_OFFSET_JAX"):
        sigma_B = getattr(B, "sigma_B", None)
        if sigma_B is None:
            raise VardynError("Basis_offset needs BASIS.sigma_B")
        return Basis_offset(State, sigma_B, name_mod_var=getattr(B, "name_mod_var", "ssh"))
    raise VardynError(f"{sup} not implemented")


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
