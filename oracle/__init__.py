"""CPU oracle for the QG 4D-Var hot path -- TEST INFRASTRUCTURE ONLY.

This package is a plain FP64 NumPy/SciPy restatement of the reference
algorithm (datlas-ocean/VarDyn, ``mapping/models/model_qg1l/jqgm.py``,
``mapping/src/mod.py``, ``mapping/src/obsop.py``,
``mapping/src/tools_4Dvar.py``, ``mapping/src/inv.py``).  It is the checker
that the HIP path is compared against.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or as the timed CPU
baseline.  The product package ``vardyn_amd`` never imports it and has no CPU
fallback.

Pinning status of each function (what independent check fixes it) is listed
in the header of each module.
"""
