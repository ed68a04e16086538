"""gascoignelib_amd: MI355X-native (gfx950) GPU multigrid solve path for Gascoigne.

The solver lives in the C-ABI library built from gascoignelib_amd/csrc
(libgascoigne_hip.so, declared in include/gascoigne_hip.h).  This package
loads it through ctypes and mirrors the reference's CudaVectorInterface /
CudaCSRMatrixInterface / CudaMgInterpolatorNested / CudaMultiLevelSolver API.
"""
from .problem import Hierarchy, Level, lattice_rhs, splitmix64_uniform  # noqa: F401
