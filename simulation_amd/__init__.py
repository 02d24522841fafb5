"""MI355X-native FDTD hot path of dsarvan/simulation.

Hand-written HIP kernels for gfx950 (libfdtd_hip.so, C-ABI in include/fdtd.h)
called through ctypes, with PyTorch-ROCm for device buffers:

  simulation_amd.fd1d   1-D Ex/Hy (fd1d/pycuda/*.py): exfield, hyfield,
                        dielectric, run (all steps in one launch)
  simulation_amd.fd2d   2-D TM Dz/Ez/Hx/Hy (fd2d/pycuda/*.py): pmlparam,
                        dielectric, dfield, efield, hfield, energy, TMSolver
                        (temporally blocked fused stepping)
  simulation_amd.build  compiles the library for gfx950
"""
from ._lib import FdtdError, LIB_PATH  # noqa: F401

__version__ = "0.1.0"
