"""Test infrastructure only: the CPU FP64 oracle for the He-4 TDDFT time step.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.
"""
