// Aggregation multigrid preconditioner of the sparse solver (setup key
// "preconditioner" = "multigrid"): the role the reference's MueLu backend
// (src/Math/TrilinosMueluSparseMatrixSolver.cpp:26-35, a multigrid
// preconditioner inside a Belos Krylov solve) plays, designed for this
// build's block-coloured ILU(0) kernels.
//
// Hierarchy (host, once per sparsity pattern): the aggregates that the ILU
// ordering already forms (bricks of about blockRows rows, or breadth-first
// graph aggregates) define an unsmoothed-aggregation prolongation P, one 1 per
// row; the coarse matrix is the Galerkin product P^T A P, whose entries are
// sums of fine entries over fixed per-slot lists (deterministic).  Each coarse
// level is a nested solver with its own ordering, SELL-64 matrix and ILU(0)
// factor, and aggregates of about multigridCoarseRows rows for the next
// level, until a level has at most multigridDenseRows rows: that one is
// inverted densely.
//
// One preconditioner apply z = B r at a level (pre- and post-smoothing by the
// level's ILU(0), M):
//   z = M^-1 r;  r_c = P^T (r - A z);  e_c = coarse solve(r_c);  z += P e_c;
//   z += M^-1 (r - A z)
// The coarse solve is the next level's apply ("V" cycle) or two flexible CG
// iterations preconditioned by it ("K" cycle, Notay & Vassilevski 2008; by
// default on the first two coarse levels, V cycles below); the last level
// multiplies by its dense inverse.  On the coarse levels M is by default the
// ILU(0) of the block-diagonal part, one aggregate per block (one launch per
// apply: on these small levels every launch costs more than its arithmetic,
// and a coloured sweep takes five).  The K cycle makes B
// nonlinear, so the outer CG runs in its flexible form (Notay 2000: beta =
// -(z, A p_old) / (p_old, A p_old)).
#pragma once

#include <vector>

#include "phase_internal.h"

struct phase_hip_solver;

namespace phase {

struct Multilevel;

// Frees the hierarchy (a new pattern, or the solver going away).
void mlFree(phase_hip_solver *s);
// The values changed: coarse matrices and factors are rebuilt before the next apply.
void mlValuesChanged(phase_hip_solver *s);
// Builds the hierarchy of the current pattern when there is none.
void mlEnsureHierarchy(phase_hip_solver *s);
// Coarse matrices, their ILU(0) factors (FP32 copies when `single`) and the
// dense inverse for the current values.
void mlEnsureValues(phase_hip_solver *s, bool single);
// Rest of a preconditioner apply after the fine pre-smoothing: on entry `out`
// holds M^-1 rhs (the fused Krylov prologue's sweep wrote it), on exit B rhs.
// Every launch returns at once when sc->done is set.
template <class T>
void mlFinishApply(phase_hip_solver *s, const Scalars *sc, const T *rhs, T *out);
// Level sizes (fine first) for reporting.
std::vector<int> mlLevelSizes(const phase_hip_solver *s);
// One launch of the fine level's residual kernel t = b - A x (FP64 matrix
// values, vectors in T), as each multigrid apply runs it; for kernel timing.
template <class T>
void mlResidLaunch(phase_hip_solver *s, const T *x, const T *b, T *t);
// The densely inverted last level: its matrix and inverse, n x n row-major
// (n = the last level size), copied to host memory.
void mlCoarsest(phase_hip_solver *s, int n, double *matrix, double *inverse);

}  // namespace phase
