"""CPU oracle (test infrastructure only): ctypes bindings to oracle/liboracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, as the checker.  The HIP product never imports it.
"""
from .pyoracle import Oracle, OracleParams, load_library, pbc  # noqa: F401
