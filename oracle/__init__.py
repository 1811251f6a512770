"""Oracles of the shallow-water step. TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use this package, and only as the checker or the timed CPU
baseline, never on the product path.

- ``cpu_ref.cpp`` / ``cpu_ref.py``: FP32 C++/OpenMP restatement of the
  reference's resident kernels, kernel by kernel (bitwise comparator and CPU
  baseline);
- ``ref64.py``: FP64 NumPy restatement of the same scheme (tolerance
  comparator, adjoint reference).
"""
