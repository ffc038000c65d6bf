"""ddhdg_amd -- MI355X (gfx950) implementation of ddhdg's per-cell HDG hot path.

Layers:
  include/ddhdg_hip.h, csrc/      C ABI + HIP kernels (libddhdg_hip.so)
  hip_lib                         ctypes binding of the C ABI (device pointers)
  problem, fe, np_solver          host mirror of the reference NPSolver/Problem API
  configs                         the diode workloads
"""
from .hip_lib import HdgContext, HdgError, LIB_PATH  # noqa: F401
from .problem import (Adimensionalizer, BoundaryConditionHandler, DIRICHLET,  # noqa: F401
                      FixedTauNPSolverParameters, HomogeneousMobility,
                      HomogeneousPermittivity, HomogeneousProblem, LinearRecombinationTerm,
                      NEUMANN, NonlinearSolverParameters, ShockleyReadHallFixedTemperature,
                      V, n, p)
from .fe import BoxMesh  # noqa: F401
