"""MI355X (gfx950) per-timestep hot path of the ssmdLBM two-phase jet solver.

The product is libssmd.so (C ABI in include/ssmd.h) built from csrc/.
This package holds its ctypes binding (abi), the parameter blocks of the
reference case and the jet workloads (cases), and a host driver that owns
device memory through torch and steps the solver through the C ABI
(solver).
"""
from . import abi, cases  # noqa: F401

__all__ = ["abi", "cases"]
