"""cubedsphere: MI355X-native dimension-split finite-volume advection on the cubed sphere.

Host side (this package): FP64 grid / metric / duo-grid tables, problem definitions,
par-file readers and the ctypes binding of libcsadv.so.  Device side: the HIP kernels
of libcsadv.so (csrc/).  The step has no CPU implementation in this package.
"""
from . import cases, config, grid, topology  # noqa: F401
from ._lib import CsadvError, load  # noqa: F401

__all__ = ["cases", "config", "grid", "topology", "CsadvError", "load"]
