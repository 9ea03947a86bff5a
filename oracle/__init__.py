"""The oracle: a CPU restatement of CORAL's BLAS algorithms (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and
only as the checker or the timed CPU baseline; libcoralhip never does.

- ``oracle.port``  -- ctypes front end of coral_port.c (CORAL's blocking, packing,
                      accumulation order and quick-return rules restated in C)
- ``oracle.ref64`` -- float64/complex128 (and exact-rational) reference formulas
"""
