"""ps3d_amd -- MI355X-native pseudo-spectral time step of ps3d (FP64 on gfx950).

The compute path is the HIP library ps3d_amd/_lib/libps3d_gpu.so behind the C ABI
of include/ps3d_gpu.h; this package is its host side (parameters, initial
conditions, ctypes binding).  Nothing here computes the solution on the CPU.
"""
from .params import Params, make_params, beltrami_box  # noqa: F401

__all__ = ["Params", "make_params", "beltrami_box", "gpu", "ics"]
