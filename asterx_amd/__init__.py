"""MI355X-native GRMHD method-of-lines substep (AsterX hot path).

The compute path is libasterx_hip.so (HIP kernels for gfx950 behind the C
ABI in include/asterx.h); this package holds its ctypes mirror (abi), patch
storage (patch), the single-patch driver (driver) and input recipes (seeds,
torus).
"""
from . import abi  # noqa: F401
