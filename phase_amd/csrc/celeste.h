// The CELESTE surface-tension model on the device (Celeste.cpp,
// CelesteStencil.cpp, SurfaceTensionForce.cpp, SurfaceTensionForceSmoothingKernel.cpp
// under src/2D/Unstructured/FiniteVolume/Multiphase/): per-mesh stencils and
// smoothing kernels built on the host once, per-step kernels over HBM fields.
#pragma once

#include "mesh.h"

namespace phase {

struct CelesteModel {
  int N = 0, F = 0;
  double sigma = 0.;
  DBuf<int> kptr, kcol;     // smoothing kernel: cells strictly inside the circle, ascending id
  DBuf<double> kw, A;       // K(x_k - x) per entry; 1 / sum K V per cell
  DBuf<int> sptr, sidx;     // stencil rows: cell id, or -(face id) - 1
  DBuf<double> ss, pw, pu;  // |r|^2 per row; rows 3, 4 of pinv (weighted, unweighted), 2 per row
  DBuf<int> dptr, dcol;     // diagonal (node) neighbours
  DBuf<double> cosT, sinT;  // contact angle per face
  DBuf<double> gt, gtf, ggt, nC, nF, kC, kF;
};

// Builds the stencils (Celeste::computeStencils) and smoothing kernels
// (SurfaceTensionForce ctor); contactAngle per face in radians (NULL: pi/2).
void celesteSetup(phase_hip_mesh *m, double sigma, double kernelWidth, int kernelType, const double *contactAngle,
                  CelesteModel &cm);
// Celeste::computeFaceInterfaceForces(gamma, gradGamma): fst on every face.
void celesteFaceForces(phase_hip_mesh *m, CelesteModel &cm, const double *gammaCell, const double *gradGammaFace,
                       double *fstFace);

}  // namespace phase
