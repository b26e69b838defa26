"""mgkmc_cpp_amd -- MI355X-native kinetic Monte Carlo + FFT
Lippmann-Schwinger hot path of mgkmc_cpp behind the reference's interface.

Native pieces (built in-tree by ``python -m mgkmc_cpp_amd.build``):
  lib/libmgkmc_gpu.so   HIP kernels for gfx950 + the C ABI (include/mgkmc_gpu.h)
  lib/libmgkmc_host.so  drop-in KmcSimulation2D/3D host driver (include/mgkmc_host.h)
"""
from . import abi  # noqa: F401
from ._lib import MgkError, gpu, host  # noqa: F401
from .solver import Context, Simulation, SimulationBatch, spectral_solver_2d, spectral_solver_3d  # noqa: F401

__all__ = ["abi", "Context", "Simulation", "SimulationBatch", "spectral_solver_2d", "spectral_solver_3d", "MgkError",
           "gpu", "host"]
