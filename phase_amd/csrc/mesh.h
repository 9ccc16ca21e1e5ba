// Device-resident unstructured 2D mesh shared by the assembly, operator and
// field kernels: topology in the reference's first-encounter face order
// (FiniteVolumeGrid2D.cpp:20-35, 84-114, 396-418) and geometry computed on the
// device (Cell.cpp:8-29, Face.cpp:9-18, CellLink.cpp:8, InteriorLink.cpp:9,
// BoundaryLink.cpp:8).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "phase_internal.h"

struct phase_hip_mesh {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int nNodes = 0, nCells = 0, nFaces = 0, nInterior = 0, nBoundary = 0;
  int maxDeg = 0;  // most interior links of one cell
  std::vector<int> fNode1, fNode2, fL, fR;
  std::vector<int> iptr, iface, inb, bptr, bface;
  phase::DBuf<double> dXY, dVol, dCC, dFC, dSf;  // vectors interleaved (x, y)
  phase::DBuf<int> dCptr, dCind, dFNode1, dFNode2, dFL, dFR;
  phase::DBuf<int> dIptr, dIface, dInb, dBptr, dBface;
  phase::DBuf<double> dIDot, dIMs, dBDot, dBMs;  // per link: dot(r, Sf_out), |r|^2
};

namespace phase {

struct Dv2 {
  double x, y;
};

// Face::outwardNorm(point), Face.cpp:48-50: dot(centroid - point, normal) > 0
// ? normal : -normal
__device__ __forceinline__ Dv2 outward(Dv2 fc, Dv2 sf, Dv2 p) {
  const double dx = fc.x - p.x, dy = fc.y - p.y;
  return (dx * sf.x + dy * sf.y) > 0. ? sf : Dv2{-sf.x, -sf.y};
}

}  // namespace phase
