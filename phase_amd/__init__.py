"""phase_amd: MI355X-native discretise-and-solve path of the Phase FV solver.

Python front-end over the C ABI in include/phase_hip.h (libphase_hip.so,
hand-written gfx950 HIP kernels).  The classes mirror the reference's
interfaces:

* HipSparseMatrixSolver -- SparseMatrixSolver (src/Math/SparseMatrixSolver.h:11-62):
  setup / setRank / set(rowPtr, colInd, vals) / setRhs / setGuess / solve / x /
  nIters / error.
* HipMesh -- FiniteVolumeGrid2D::init(nodes, cptr, cind)
  (src/2D/Unstructured/FiniteVolumeGrid2D/FiniteVolumeGrid2D.cpp:20-35).
* HipSparseMatrixSolver.assemble_laplacian -- fv::laplacian
  (src/2D/Unstructured/FiniteVolume/Discretization/Laplacian.h:7-167) assembled
  on the device into the solver's system.
* HipEquation -- FiniteVolumeEquation<T> / CrsEquation on the device: the fv::
  operators (laplacian, ddt, div), += / -= / == / *= / relax, export in the
  reference's layout, and HipSparseMatrixSolver.set_equation for the solve.
* HipMesh.src_div / src_src / grad / interpolate_faces -- src::div, src::src,
  ScalarGradient, interpolateFaces on device fields.
* HipFractionalStep -- FractionalStep::solve, the reference's projection solver
  (src/2D/Unstructured/Solvers/FractionalStep.cpp:36-117), over the above.
* HipMultiphase -- FractionalStepMultiphase::solve (CICSAM volume fraction,
  properties, gravity source, CELESTE surface tension, momentum, pressure)
  (src/2D/Unstructured/Solvers/FractionalStepMultiphase.cpp:58-218).
* HipGhostCellIB -- GhostCellImmersedBoundary for circular bodies (cell types,
  bcs / velocityBcs rows, applyHydrodynamicForce); with
  HipFractionalStep.immersed_boundary it gives FractionalStepGCIB::solve.
* HipSimple -- a SIMPLE outer loop over the above (the README's uEqn/pCorrEqn).

There is no CPU fallback: loading fails loudly when the library is absent.
"""
from .hip import (BC_FIXED, BC_NORMAL_GRADIENT, BC_OUTFLOW, BC_PARTIAL_SLIP, BC_SYMMETRY,
                  MODE_FP64, MODE_LEARNED, MODE_MIXED, DeviceArray, HipEquation, HipFractionalStep,
                  HipGhostCellIB, HipMesh, HipMultiphase, HipSimple, HipSparseMatrixSolver, PhaseHipError,
                  device_count, library_path, load_library)

__all__ = [
    "BC_FIXED", "BC_NORMAL_GRADIENT", "BC_OUTFLOW", "BC_PARTIAL_SLIP", "BC_SYMMETRY",
    "MODE_FP64", "MODE_LEARNED", "MODE_MIXED", "DeviceArray", "HipEquation", "HipFractionalStep",
    "HipGhostCellIB", "HipMultiphase", "HipMesh", "HipSimple", "HipSparseMatrixSolver",
    "PhaseHipError", "device_count", "library_path", "load_library",
]
