// The solver handle behind the phase_hip_solver_* C entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "phase_internal.h"

namespace phase {

// Largest GMRES restart length (Krylov basis vectors per cycle).
constexpr int kGmMaxRestart = 64;

template <class T>
struct KrylovVecs {
  DBuf<T> r, r0, p, v, s, t, ph, sh, y, x;
  void alloc(int n) {
    for (DBuf<T> *b : {&r, &r0, &p, &v, &s, &t, &ph, &sh, &y, &x}) b->alloc(n);
  }
};

struct MeshBinding;  // mesh.hip: link -> slot maps of an assembled operator
struct GnnState;     // gnn.hip: learned correction network
struct Multilevel;   // multilevel.hip: aggregation multigrid hierarchy

}  // namespace phase

struct phase_hip_solver {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;

  // setup() keys
  double tol = 1e-8;
  int maxIters = 500;
  int mode = PHASE_HIP_MODE_FP64;
  double innerTol = 0.1;
  int blockRows = 32;  // rows per ILU aggregate; blocks pack same-colour aggregates up to 64 rows
  double fallbackRatio = 0.5;
  int maxOuter = 100;
  int chunk = 8;  // iterations enqueued between host checks
  int geometric = 1; // cell centroids known: brick aggregates in three colours (else graph aggregates)
  int cg = 0;        // Krylov method: 0 = BiCGStab, 1 = preconditioned CG (symmetric systems)
  int gmres = 0;     // 1: flexible restarted GMRES (gmres.hip)
  int gmresRestart = 30;
  phase::DBuf<double> gmV, gmZ, gmH, gmSmall, gmPart;  // GMRES basis, preconditioned basis, Hessenberg, scalars
  phase::DBuf<float> gmV32, gmZ32;                    // the same two in mode mixed
  int unionPattern = 0;  // keep the stored pattern when a new one lies inside it (its other entries 0)
  int blockJacobi = 0;   // ILU(0) of the block-diagonal part only, one colour (coarse multigrid levels)
  int iluLimits = 1;     // 0: the ordering's ILU(0) is never applied (Jacobi multigrid levels), any row width
  int precond = 0;       // 0 = ILU(0), 1 = aggregation multigrid with ILU(0) smoothing
  int mlKCycleSetting = -1;  // multigrid coarse solves: 1 = K cycle, 0 = V cycle, -1 = K for CG, V for BiCGStab
  int mlKCycle = 0;          // the choice in force for the current solve
  int mlCoarseRows = 16; // rows per aggregate below the finest level
  int mlKLevels = 2;     // coarse levels (from the first) whose solves are K cycles; V cycles below
  int mlFinePost = 1;    // post-smoothing on the finest level (0: pre-smoothing only)
  int mlBlockJacobiFrom = 0;   // coarse levels from this one (0 = the first) smooth with block-Jacobi ILU(0)
  int mlDenseRows = 256; // a level of at most this many rows is inverted densely
  int mlJacobiFrom = -1;  // coarse levels from this one (0 = all) smooth with damped Jacobi instead of ILU(0);
                          // -1: all with CG, none with BiCGStab
  double mlJacobiWeight = 0.8;  // the damping weight (0.8: fewer iterations than 2/3 on every benchmarked system)
  int mlDensePivot = 0;  // 1: the coarsest inverse always pivots (0: only when a pivot is small)

  // pattern and ordering
  int rank = 0;
  bool havePattern = false;
  std::vector<double> cellXY;  // row centroids of the pattern (brick aggregates), empty when unknown
  std::vector<int> patRowPtr, patColInd;
  phase::Ordering ord;

  phase::DBuf<phase::BlockInfo> dBlocks;
  phase::DBuf<phase::RowInfo> dRows;
  phase::DBuf<phase::LaneRec> dLanes;
  phase::DBuf<int> dLCol, dUCol, dUpdPtr, dUpdSrc, dUpdDst;
  phase::DBuf<int> dSliceOff, dSellCol, dSellSrc, dFacSrc, dPerm;

  // values: caller slots (FP64), SpMV copies, ILU factors
  phase::DBuf<double> dVals;
  phase::DBuf<double> dSell64, dFac64;
  phase::DBuf<float> dSell32, dFac32;
  bool haveValues = false, sell32Valid = false, fac64Valid = false, fac32Valid = false;

  // right-hand side / solution, solver order
  phase::DBuf<double> dB, dX, dR64, dXBest;
  bool haveRhs = false, haveGuess = false, haveSolution = false;
  phase::KrylovVecs<double> k64;
  phase::KrylovVecs<float> k32;

  // synchronisation and reductions
  phase::DBuf<int> dErr;  // bit 2: zero pivot in the factorisation; 4: singular coarsest multigrid level;
                          // 8: small pivot in its unpivoted inverse (inverted again with pivoting)
  phase::DBuf<double> dPartials;
  phase::DBuf<unsigned> dCounter;
  phase::DBuf<phase::Scalars> dSc;  // [0] inner run, [1] outer residual, [2] misc
  phase::Scalars *hSc = nullptr;    // pinned mirror of dSc
  phase::DBuf<double> dMisc;        // [0] ||b||^2

  // results of the last solve
  int nIters = 0, innerIters = 0, outerPasses = 0, fallbacks = 0, learnedSteps = 0;
  double error = 0.0, solveMs = 0.0;

  // learned correction (gnn.hip) and its FP64 work vectors
  phase::GnnState *gnn = nullptr;
  phase::DBuf<double> dE64, dQ64;
  int learnedObjective = 1;  // training loss: 0 the error's direction, 1 the residual the step leaves
  int lastCaptureIter = 0;

  // pattern of the last equation handed over by phase_hip_solver_set_equation
  phase::DBuf<int> dEqCols, dEqFlag;
  int eqStride = 0;

  // multigrid hierarchy of the current pattern (multilevel.hip)
  phase::Multilevel *ml = nullptr;

  // mesh-assembled operator (mesh.hip)
  phase::MeshBinding *binding = nullptr;
  phase_hip_mesh *boundMesh = nullptr;
  int lastVariant = -1;
};

namespace phase {
// Upload a new pattern (rebuilding the ordering) -- used by set() and by the
// mesh assembly path.
void solverSetPattern(phase_hip_solver *s, int n, const int *rowPtr, const int *colInd,
                      const double *cellXY = nullptr);
// dVals changed on the device: refresh SpMV copies, invalidate factors.
void solverValuesChanged(phase_hip_solver *s);
void solverFreeBinding(phase_hip_solver *s);
// Raise a boundary-type error flagged by the last device assembly, if any.
void checkAssembly(phase_hip_solver *s);
// FP32 copies of A and the FP32 ILU(0) factor are current.
void solverEnsureSingle(phase_hip_solver *s);
// out = U^{-1} L^{-1} in with the ILU(0) factor fac (solver order; tmp holds
// the forward-sweep result).  Instantiated for float and double.
template <class T>
void iluApplyDev(phase_hip_solver *s, T *fac, const T *in, T *tmp, T *out);
// Flexible restarted GMRES from the current x (gmres.hip), FP64 or FP32 cycles
// restarted from the FP64 residual; b2 = ||b||^2.
void solveGmres(phase_hip_solver *s, double b2);
}  // namespace phase
