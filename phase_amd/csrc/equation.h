// Device-resident FiniteVolumeEquation<T> / CrsEquation (see equation.hip):
// ncomp * nCells rows in the IndexMap block layout, each row up to `stride`
// (column, value) pairs in first-touch order, plus the stored rhs
// (A x + rhs = 0).
#pragma once

#include <string>

#include "mesh.h"
#include "phase_internal.h"

struct phase_hip_equation {
  phase_hip_mesh *mesh = nullptr;
  int ncomp = 1, rank = 0, stride = 0;
  phase::DBuf<int> cols, cnt;
  phase::DBuf<double> vals, rhs;
  phase::DBuf<int> dErr;
  bool compact = false;  // rows rebuilt by += / -= (no empty slots)
  std::string err;
};
