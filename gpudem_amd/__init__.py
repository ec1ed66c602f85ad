"""gdem: MI355X-native DEM contact engine (drop-in for gpudem's per-step path).

The compute path is libgdem.so (hand-written gfx950 HIP kernels behind the C
ABI in include/gdem.h); this package holds its ctypes binding, scene
producers and the build script.
"""
from . import scenes  # noqa: F401
from .engine import (GDEM_FP64, GDEM_MIXED, Batch, Context, GdemError, Scene, bbox_rule,  # noqa: F401
                     lib, material)

__all__ = ["Context", "Scene", "Batch", "GdemError", "GDEM_FP64", "GDEM_MIXED", "material", "bbox_rule",
           "lib", "scenes"]
