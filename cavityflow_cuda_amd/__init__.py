"""MI355X-native D3Q19 lattice-Boltzmann alpha/beta A-A step.

A from-scratch HIP (gfx950) implementation of the hot path of
saleemayman/cavityflow_cuda behind that project's CLbmSolver operator surface:

* ``include/lbm_abi.h`` / ``libcavityflow_hip.so``: the C ABI (create, alpha and
  beta steps over boxes, population/flag/velocity/density box copies, face
  pack/unpack, FP64 reductions);
* ``solver.LbmSolverHIP``: the CLbmSolver<T> method surface over that ABI;
* ``controller.CavityController``: the CController step driver (step parity,
  lid set-up, on-device face exchange between subdomains, benchmark figures).

There is no CPU fallback: using the package without the built HIP library
raises.
"""
from ._lib import (BACK, BF16S, BOTTOM, F16C, F16S, F32, F64, FLUID, FRONT, GHOST_LAYER, LEFT,
                   OBSTACLE, RIGHT, TOP, VELOCITY_INJECTION, LbmError)
from .solver import LbmSolverHIP, compute_parameters, exchange_faces, face_transfers
from .controller import CavityController, cavity_parameters

__all__ = [
    "LbmSolverHIP", "CavityController", "cavity_parameters", "compute_parameters", "LbmError",
    "exchange_faces", "face_transfers",
    "OBSTACLE", "FLUID", "VELOCITY_INJECTION", "GHOST_LAYER", "LEFT", "RIGHT", "BOTTOM", "TOP",
    "BACK", "FRONT", "F32", "F64", "F16S", "BF16S", "F16C",
]
