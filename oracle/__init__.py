"""ORACLE — test infrastructure only.

A CPU restatement of the reference's 2D/3D deterministic substep and state
hash (see sim2d_ref.cpp / sim3d_ref.cpp), loaded through ctypes.  Only
tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py
may import this package; the product (``sunaengine_amd``) never does.
"""
from .oracle import OracleWorld, OracleWorld3, build, lib_path  # noqa: F401
