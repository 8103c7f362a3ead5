"""mps_cuda_amd — MI355X (gfx950) SPH time step of MPS_CUDA behind a C ABI.

The compute lives in libmps_sph.so (hand-written HIP kernels, see csrc/). This package
is the Python handle used by the tests and the benchmark; the C++ classes mirroring the
reference's System/SPHMaterial/MeshObject/SpatialHash are in include/mps_system.hpp.
"""
from ._capi import (  # noqa: F401
    ARRAYS, KERNEL_CUBIC, KERNEL_POLY6_SPIKY, PHASE, VISC_CG, VISC_JACOBI, VISC_NONE,
    Config, HipSim, MPSError, StepStats, default_config, exported_symbols, load, read_obj,
)

__all__ = ["HipSim", "Config", "StepStats", "MPSError", "default_config", "load", "read_obj", "PHASE", "ARRAYS"]
