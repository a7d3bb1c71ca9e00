"""CPU FP64 oracle for the dnsjax time-step hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain NumPy/SciPy float64, the algorithm of the
reference solver's time-step path (``timestep.make_stepper`` and the flow
callables behind it).  Every function cites the reference file:line it
follows (paths relative to the reference's ``src/dnsjax/``).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
(``dnsjax_amd``) never imports, links or executes anything in here.

Pinning (what fixes each function independently of its own formula) is
listed in each module's header and exercised by ``tests/test_oracle_*.py``.
"""
