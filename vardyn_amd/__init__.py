"""vardyn_amd -- MI355X-native QG 4D-Var hot path (forward, adjoint, cost gradient)
with the call surface of datlas-ocean/VarDyn.

The numerics run in libvardyn_hip.so (HIP kernels for gfx950, C ABI in
include/vardyn.h).  Modules:

* ``jqgm``        -- ``Qgm`` drop-in for mapping/models/model_qg1l/jqgm.py
* ``qg``          -- ``QgPlan``: device plan + C-ABI calls on torch tensors
* ``autograd``    -- ``QgStep``: the step as a torch.autograd.Function (hand adjoint / TLM)
* ``obsop``       -- L3 observation operator (mapping/src/obsop.py)
* ``mod``, ``state``, ``basis`` -- ``Model_qg1l``, ``State``, ``Basis_offset`` / ``Basis_ic``
* ``fourdvar``    -- ``FourDVar``: device-resident (batched) cost and gradient, HIP-graph replay
* ``problem``     -- assembly of synthetic 4D-Var problems
* ``tools_4Dvar`` -- ``Cov``, ``Variational.cost_and_grad``, ``grad_test``
* ``inv``         -- checkpoint schedule and L-BFGS driver (mapping/src/inv.py)
* ``synthetic``   -- seeded synthetic problems
"""
__version__ = "0.1.0"
