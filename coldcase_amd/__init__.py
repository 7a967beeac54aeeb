"""coldcase_amd -- MI355X-native engine for the covering-code search hot path.

The compute runs in libcovhip.so (hand-written HIP kernels for gfx950) behind
a C ABI (include/covhip.h); this package is the host side:

  engine      ctypes binding of the ABI and the `Cell` handle
  verify_cov  drop-in for cov/verify_cov.py (verify_pure / verify_numpy /
              covering_radius / CLI), backed by the GPU kernels
  covsearch   drop-in for the covsearch CLI contract (RESULT line, --out file)
  batch       many independent cells in one persistent launch + autotuner
  serve       continuous-batching request loop with HBM-budget admission
"""
from .engine import Batch, Cell, CovError, Fleet, lib, lib_path  # noqa: F401

__all__ = ["Batch", "Cell", "CovError", "Fleet", "lib", "lib_path"]
