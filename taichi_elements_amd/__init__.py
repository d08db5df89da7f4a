"""MI355X-native MLS-MPM substep engine behind taichi_elements' MPMSolver API.

``from taichi_elements_amd.mpm_solver import MPMSolver`` replaces
``from engine.mpm_solver import MPMSolver``; the substep runs in the HIP
kernels of ``libmpm_hip.so`` (see include/mpm_abi.h).
"""
from .mpm_solver import MPMSolver  # noqa: F401
from .batch import MPMSolverBatch  # noqa: F401

__all__ = ["MPMSolver", "MPMSolverBatch"]
