/*
 * phase_hip.h -- C ABI of the MI355X (gfx950) discretise-and-solve library.
 *
 * It sits behind two interfaces of the reference (obrienadam/Phase; paths are
 * relative to its root):
 *
 *  1. the sparse-solver plugin class SparseMatrixSolver
 *     (src/Math/SparseMatrixSolver.h:11-62), selected by the string
 *     LinearAlgebra.<eqn>.lib through SparseMatrixSolverFactory::create
 *     (src/Math/SparseMatrixSolverFactory.cpp:26-41).  A C++ adapter class
 *     (phase_amd/host/HipSparseMatrixSolver.h) implements that class on top of
 *     the phase_hip_* solver entry points below; its factory name is "hip".
 *
 *  2. the fv:: operator API (src/2D/Unstructured/FiniteVolume/Discretization/
 *     Laplacian.h:7-167) over a FiniteVolumeGrid2D built by
 *     FiniteVolumeGrid2D::init(nodes, cptr, cind)
 *     (src/2D/Unstructured/FiniteVolumeGrid2D/FiniteVolumeGrid2D.cpp:20-35).
 *     phase_hip_mesh_create takes exactly those arrays; the assembly entry
 *     points build the equation's CSR and right-hand side on the device.
 *
 * Conventions (the reference's, unless stated):
 *  - Index = int32, Scalar = double (src/Types/Types.h:1-14).
 *  - An equation stores A and rhs with A x + rhs = 0; phase_hip_solve solves
 *    A x = -rhs exactly as FiniteVolumeEquation<T>::solve passes -rhs_ to the
 *    solver (src/2D/Unstructured/FiniteVolume/Equation/FiniteVolumeEquation.tpp:73).
 *    phase_hip_set_rhs, like SparseMatrixSolver::setRhs, takes the right-hand
 *    side b of A x = b directly.
 *  - Host arrays are caller-owned and copied; device memory is owned by the
 *    handle.  A handle is not thread-safe; each has its own HIP stream.
 *  - Every function returns 0 on success or a negative PHASE_HIP_E* code; the
 *    message is available from phase_hip_last_error(handle) (or, for a failed
 *    create, phase_hip_last_error(NULL)).  The C++ adapter turns it into the
 *    reference's Exception(class, method, description) (src/System/Exception.h).
 *  - Boundary types use the values of FiniteVolumeField<T>::BoundaryType
 *    (src/2D/Unstructured/FiniteVolume/Field/FiniteVolumeField.h:12).
 */
#ifndef PHASE_HIP_H
#define PHASE_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phase_hip_solver phase_hip_solver;
typedef struct phase_hip_mesh phase_hip_mesh;

enum {
  PHASE_HIP_OK = 0,
  PHASE_HIP_EINVAL = -1,     /* bad argument or inconsistent sizes */
  PHASE_HIP_EDEVICE = -2,    /* HIP runtime error */
  PHASE_HIP_ESTATE = -3,     /* call out of order (no pattern, no mesh ...) */
  PHASE_HIP_ENOCONV = -4,    /* solve stopped at maxIters above tolerance */
  PHASE_HIP_EBREAKDOWN = -5, /* zero pivot or Krylov breakdown */
  PHASE_HIP_EBOUNDARY = -6   /* unsupported boundary type for the operator */
};

enum {
  PHASE_HIP_BC_FIXED = 0,
  PHASE_HIP_BC_NORMAL_GRADIENT = 1,
  PHASE_HIP_BC_SYMMETRY = 2,
  PHASE_HIP_BC_OUTFLOW = 3,
  PHASE_HIP_BC_PARTIAL_SLIP = 4
};

/* Smoothing kernels of the surface-tension model
 * (SurfaceTensionForce::SmoothingKernel::Type, SurfaceTensionForce.h). */
enum { PHASE_HIP_KERNEL_PESKIN = 0, PHASE_HIP_KERNEL_POW6 = 1, PHASE_HIP_KERNEL_POW8 = 2 };

/* Solve modes (phase_hip_setup key "mode"). */
enum {
  PHASE_HIP_MODE_FP64 = 0,   /* BiCGStab + ILU(0), all FP64 */
  PHASE_HIP_MODE_MIXED = 1,  /* FP32 BiCGStab + ILU(0) inside FP64 iterative refinement */
  PHASE_HIP_MODE_LEARNED = 2 /* MIXED with the learned initial guess / flexible preconditioner */
};

const char *phase_hip_last_error(const phase_hip_solver *s);
int phase_hip_device_count(int *count);

/* Device memory for callers that keep operator inputs resident in HBM (the
 * *_device entry points).  toDevice: 1 host->device, 0 device->host. */
int phase_hip_malloc(int device, size_t bytes, void **ptr);
int phase_hip_free(void *ptr);
int phase_hip_memcpy(void *dst, const void *src, size_t bytes, int toDevice);
/* Device-to-device copy through a plain copy kernel (each byte read once and
 * written once; pointers and size 16-byte aligned, else EINVAL).  Also the
 * known-size copy against which HBM traffic counters are scaled. */
int phase_hip_copy_device(void *dst, const void *src, size_t bytes);

/* ------------------------------------------------------------------ solver */

/* SparseMatrixSolverFactory::create("hip") -- SparseMatrixSolverFactory.cpp:26-41 */
int phase_hip_solver_create(phase_hip_solver **s, int device);
int phase_hip_solver_destroy(phase_hip_solver *s);

/* SparseMatrixSolver::setup(ptree) -- TrilinosBelosSparseMatrixSolver.cpp:44-86.
 * Keys: "tolerance" (1e-8), "maxIters" (500), "solver" (BICGSTAB | CG |
 * GMRES, the reference's Belos names; GMRES is flexible and restarted, the
 * default of the reference's MueLu backend,
 * TrilinosMueluSparseMatrixSolver.cpp:47; in mode mixed or learned its
 * cycles run in FP32 from the FP64 residual, each until its estimate drops
 * 1e5-fold, continuing in FP64 when they stall; no network under GMRES),
 * "gmresRestart" (basis vectors per GMRES cycle, 1 to 64, 30), "mode"
 * (fp64|mixed|learned),
 * "innerTolerance" (mixed: residual reduction between FP64 refinement points, 0.1),
 * "maxOuter" (mixed: FP64 refinement passes, 100), "chunk" (iterations queued
 * between host checks, 8), "blockRows" (rows per ILU aggregate, <= 64,
 * default 32; aggregates of one colour are packed into blocks of up to 64
 * rows), "ordering" (geometric | graph aggregates), "pattern" (exact | union),
 * "fallbackRatio" (learned: minimum per-iteration residual reduction before
 * falling back, 0.5), "learnedObjective" (the network's training loss:
 * RESIDUAL, the default, the part of the pair's residual that the
 * minimal-residual step of the output leaves; ERROR, the direction of the
 * captured correction),
 * "preconditioner" (ILU = Ifpack2 RILUK(0) as the reference's Belos backend,
 * the default; MULTIGRID = aggregation multigrid, ILU(0) smoothing on the
 * finest level and damped Jacobi on the coarse ones, the role of the
 * reference's MueLu backend, TrilinosMueluSparseMatrixSolver.cpp:26-35),
 * "multigridCycle" (K | V | auto: K cycles with flexible CG, V cycles with
 * BiCGStab), "multigridKLevels" (coarse levels, from the first, whose solves
 * are K cycles; V cycles below; 2), "multigridJacobiFrom" (coarse levels
 * from this index on, 0 being the first, smooth with damped Jacobi, the
 * others with ILU(0); AUTO: all with CG, none with BiCGStab),
 * "multigridJacobiWeight" (the damping, 0.8),
 * "multigridBlockJacobiFrom" (coarse ILU(0) levels from this index on use
 * the ILU(0) of their block-diagonal part in one launch instead of the
 * block-coloured ILU(0); 0), "multigridFinePostSmoothing" (1 | 0),
 * "multigridCoarseRows" (rows per aggregate below the finest level, 16),
 * "multigridDenseRows" (a level of at most this many rows is inverted
 * densely, 256, at most 512), "multigridDensePivoting" (AUTO: without
 * pivoting unless a pivot is small | ALWAYS).  Unknown keys are ignored, as
 * the reference ignores ptree keys it does not read. */
int phase_hip_solver_setup(phase_hip_solver *s, const char *key, const char *value);

/* SparseMatrixSolver::setRank(int) -- SparseMatrixSolver.h:23 */
int phase_hip_solver_set_rank(phase_hip_solver *s, int rank);

/* SparseMatrixSolver::set(rowPtr, colInd, vals) -- SparseMatrixSolver.h:30-32,
 * TrilinosSparseMatrixSolver.cpp:71-96.  nslots = rowPtr[n]; colInd < 0 marks
 * an empty slot and is skipped.  When the pattern equals the previous call's
 * only the values are uploaded; otherwise the ordering (RCM, ILU blocks and
 * block colouring) is rebuilt. */
int phase_hip_solver_set(phase_hip_solver *s, int n, const int *rowPtr, const int *colInd,
                         const double *vals);

/* SparseMatrixSolver::setRhs / setGuess -- SparseMatrixSolver.h:36-38 */
int phase_hip_solver_set_rhs(phase_hip_solver *s, const double *b);
int phase_hip_solver_set_guess(phase_hip_solver *s, const double *x0);

/* setGuess(0) without a host copy: the next solve starts from x = 0. */
int phase_hip_solver_zero_guess(phase_hip_solver *s);

/* SparseMatrixSolver::solve() -- SparseMatrixSolver.h:40; returns error() via
 * *relres: the FP64 true relative residual ||b - A x||_2 / ||b||_2. */
int phase_hip_solver_solve(phase_hip_solver *s, double *relres);

/* SparseMatrixSolver::x(Index) for all rows, nIters(), error() --
 * SparseMatrixSolver.h:46-54 */
int phase_hip_solver_get_x(phase_hip_solver *s, double *x);
int phase_hip_solver_n_iters(const phase_hip_solver *s);
double phase_hip_solver_error(const phase_hip_solver *s);

/* Per-solve statistics: inner iterations, refinement passes, fallbacks taken
 * (learned mode: passes where the network step was rejected), device
 * milliseconds of the last solve (HIP events on the handle's stream). */
int phase_hip_solver_stats(const phase_hip_solver *s, int *innerIters, int *outerPasses,
                           int *fallbacks, double *solveMs);

/* Shape of the device ordering of the current pattern: ILU blocks (<= 64 rows,
 * one wavefront each) and block colours (= sweep launches per triangular
 * solve).  No counterpart in the reference. */
int phase_hip_solver_ordering_info(const phase_hip_solver *s, int *nBlocks, int *nColours);
/* Rows of each level of the multigrid hierarchy, finest first (the last level
 * is inverted densely); nLevels = 0 before the first multigrid solve of a
 * pattern.  No counterpart in the reference. */
int phase_hip_solver_multigrid_info(const phase_hip_solver *s, int *nLevels, int *sizes, int cap);
/* The densely inverted last level of the multigrid hierarchy after a solve:
 * its Galerkin matrix and the inverse the cycle applies, n x n row-major in
 * the same row order (n must equal the last entry of
 * phase_hip_solver_multigrid_info).  For checking; no counterpart in the
 * reference. */
int phase_hip_solver_multigrid_coarsest(phase_hip_solver *s, int n, double *matrix, double *inverse);

/* Export the system held on the device, in the ORIGINAL row/column numbering
 * and compact CSR form (empty slots dropped), each row in the solver's slot
 * order.  nnz is queried with rowPtr == NULL.  Used for parity checks. */
int phase_hip_solver_export(phase_hip_solver *s, int *nnz, int *rowPtr, int *colInd,
                            double *vals, double *rhs);

/* Device SpMV y = A x on the held system (original numbering in and out).
 * FP64 when single == 0, FP32 values and vectors otherwise. */
int phase_hip_solver_spmv(phase_hip_solver *s, int single, const double *x, double *y);

/* phase_hip_solver_spmv without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * operand and product are staged in a fixed 32 MiB device array that the
 * library holds on each GPU it runs on; ranks above 2,097,152 rows are
 * rejected (PHASE_HIP_EINVAL), and single != 0 needs the FP32 copy of the
 * matrix that a mixed solve or phase_hip_solver_spmv(single = 1) builds
 * (PHASE_HIP_ESTATE otherwise).  Every call on one GPU shares that array:
 * calls that overlap on one GPU (several host threads) race. */
int phase_hip_solver_spmv_noalloc(phase_hip_solver *s, int single, const double *x, double *y);

/* The ILU(0) factor computed on the device, in SOLVER order (rows and columns
 * permuted), compact CSR with sorted columns: L strictly below the diagonal
 * (unit diagonal implied), U on and above.  perm[k] = original row of solver
 * row k.  nnz is queried with rowPtr == NULL. */
int phase_hip_solver_ilu_export(phase_hip_solver *s, int *nnz, int *perm, int *rowPtr,
                                int *colInd, double *vals);

/* Timing hook for the benchmark: runs each kernel `reps` times on the
 * resident system and writes the mean device time (ms, HIP events on the
 * handle's stream) to ms[0..nms): 0 SpMV FP64, 1 SpMV FP32, 2 ILU forward
 * sweep FP64 (all colours), 3 backward sweep FP64, 4 forward FP32,
 * 5 backward FP32, 6 ILU(0) factorisation FP64 (value gather + factor),
 * 7 factorisation FP32, 8 one backward-sweep launch (kUSolve, FP64; mean over
 * the colours a solve launches it for), 9 the fused last-colour launch
 * (kLUSolve, FP64), 10 one forward-sweep launch (kLSolve, FP64, mean as 8),
 * 11 / 12 the FP64 / FP32 SpMV cache-cold (a 1 GiB copy before each launch
 * evicts L2 and the last-level cache; the events bracket the one launch),
 * 13..15 as 8..10 with the FP32 factor (mixed mode's sweeps), 16 / 17 the
 * forward-sweep and fused launches in the form CG runs them (the prologue
 * also moves x += alpha p and r -= alpha q), FP64, 18 / 19 the same in FP32,
 * 20 / 21 the multigrid cycle's fine-level residual t = b - A x (FP64 matrix
 * values; FP32 / FP64 vectors).
 * Benchmark support; no counterpart in the reference. */
int phase_hip_solver_time_kernels(phase_hip_solver *s, int reps, double *ms, int nms);

/* Algorithmic bytes of the entries of phase_hip_solver_time_kernels: every
 * matrix or factor entry touched (value + column index), every vector element
 * read or written once, one gathered value per off-block factor entry, and the
 * block / lane records the sweeps read.  Benchmark support. */
int phase_hip_solver_kernel_bytes(phase_hip_solver *s, double *bytes, int nb);

/* ------------------------------------------------------------------ learned part
 * No counterpart in the reference.  A message-passing network on the matrix
 * graph (6 layers, 64 hidden, BF16 MFMA / FP32 accumulate) trained on
 * (residual, correction) pairs captured from the solver's own FP64 solves; in
 * mode "learned" each FP64 refinement pass first applies the network's
 * correction as a minimal-residual step and keeps it only when it reduces the
 * residual by at least "fallbackRatio" (0.5), otherwise the solve continues
 * with plain ILU(0) refinement.  Requires a symmetric matrix. */

/* Capture a pair every `everyIters` iterations of FP64 solves (0 = off), up to
 * maxPairs pairs in device memory (2 * maxPairs * rank floats). */
int phase_hip_learned_capture(phase_hip_solver *s, int everyIters, int maxPairs);
int phase_hip_learned_n_pairs(const phase_hip_solver *s, int *count);
/* Checkpoint of the captured pairs: pairs [first, first + count) copied out
 * in original numbering, pair-major (r and e: count * rank floats each). */
int phase_hip_learned_pairs(phase_hip_solver *s, int first, int count, float *r, float *e);
/* Append count pairs in that layout to the store phase_hip_learned_capture set
 * up (EINVAL beyond its maxPairs; ESTATE without a store). */
int phase_hip_learned_add_pairs(phase_hip_solver *s, int count, const float *r, const float *e);
/* `steps` Adam steps (one captured pair each) at learning rate lr; *lastLoss is
 * the mean loss over those steps.  Loss = 1 - cos^2(network output, target
 * correction): only the direction matters, the solver scales the step. */
int phase_hip_learned_train(phase_hip_solver *s, int steps, double lr, int seed, double *lastLoss);
/* Flat FP32 parameters (checkpoint); count queried with params == NULL. */
int phase_hip_learned_params(phase_hip_solver *s, float *params, int *count);
int phase_hip_learned_set_params(phase_hip_solver *s, const float *params, int count);
/* Forward + backward for residual r, ILU-preconditioned residual z and target
 * (in network output units, original numbering): output, loss and flat
 * gradient. */
int phase_hip_learned_eval(phase_hip_solver *s, const double *r, const double *z, const double *target,
                           double *out, double *loss, float *grad);
/* e = network correction of residual r (original numbering). */
int phase_hip_learned_apply(phase_hip_solver *s, const double *r, double *e);

/* ------------------------------------------------------------------ mesh */

/* FiniteVolumeGrid2D::init(nodes, cptr, cind) -- FiniteVolumeGrid2D.cpp:20-35,
 * createCell :84-114, init() :396-418.  xy: 2*nNodes doubles; cell c has nodes
 * cind[cptr[c]..cptr[c+1]) counter-clockwise. */
int phase_hip_mesh_create(phase_hip_mesh **m, int device, int nNodes, const double *xy,
                          int nCells, const int *cptr, const int *cind);
int phase_hip_mesh_destroy(phase_hip_mesh *m);
const char *phase_hip_mesh_last_error(const phase_hip_mesh *m);
int phase_hip_mesh_sizes(const phase_hip_mesh *m, int *nCells, int *nFaces,
                         int *nInteriorFaces, int *nBoundaryFaces);
/* Face topology (face ids in the reference's first-encounter order). */
int phase_hip_mesh_faces(const phase_hip_mesh *m, int *lNode, int *rNode, int *lCell,
                         int *rCell);
/* Geometry computed on the device: cell volume / centroid (Cell.cpp:8-29),
 * face centroid / unnormalised normal (Face.cpp:9-18). */
int phase_hip_mesh_geometry(const phase_hip_mesh *m, double *cellVolume,
                            double *cellCentroid, double *faceCentroid, double *faceSf);

/* ------------------------------------------------------------------ operators */

/* fv::laplacian into the solver's system (pattern = the mesh's cell graph):
 *   variant 0: laplacian(Scalar gamma, phi, theta)  Laplacian.h:7-47
 *   variant 1: laplacian(Scalar gamma, phi)         Laplacian.h:49-83
 *   variant 2: laplacian(field gamma, phi, theta)   Laplacian.h:85-130
 *   variant 3: laplacian(field gamma, phi)          Laplacian.h:132-167
 * gammaFace / gamma0Face: per face (variants 2,3); bcType: per face (interior
 * entries ignored; NULL = NORMAL_GRADIENT everywhere, the reference default,
 * FiniteVolumeField.tpp:104-109); phiFace: boundary values; phi0Cell/phi0Face:
 * previous time level (theta variants).  Pointers may be NULL when unused.
 * The first call binds the solver to the mesh (its pattern); later calls
 * replace values and rhs in place. */
int phase_hip_assemble_laplacian(phase_hip_solver *s, phase_hip_mesh *m, int variant,
                                 double gamma, const double *gammaFace,
                                 const double *gamma0Face, double theta, const int *bcType,
                                 const double *phiFace, const double *phi0Cell,
                                 const double *phi0Face);

/* As phase_hip_assemble_laplacian, with every array argument a DEVICE pointer
 * (e.g. from phase_hip_malloc) so operator inputs stay resident in HBM.  The
 * boundary-type check then runs in the kernel and an unsupported type is
 * reported (PHASE_HIP_EBOUNDARY) by the next solve or export. */
int phase_hip_assemble_laplacian_device(phase_hip_solver *s, phase_hip_mesh *m, int variant,
                                        double gamma, const double *gammaFace,
                                        const double *gamma0Face, double theta,
                                        const int *bcType, const double *phiFace,
                                        const double *phi0Cell, const double *phi0Face);

/* CrsEquation::operator==(Vector) / operator-=(Vector) -- CrsEquation.cpp:282-311:
 * rhs -= v (v per row, original numbering).  _device: v is a device pointer. */
int phase_hip_equation_sub_rhs(phase_hip_solver *s, const double *v);
int phase_hip_equation_sub_rhs_device(phase_hip_solver *s, const double *v);

/* The held equation in the reference's padded CrsEquation layout
 * (CrsEquation.cpp:9-14: 5 slots per row, -1 = empty, first-touch column order
 * of the last assembled variant) with the stored rhs (A x + rhs = 0).
 * nslots is queried with rowPtr == NULL. */
int phase_hip_equation_export_padded(phase_hip_solver *s, int *nslots, int *rowPtr,
                                     int *colInd, double *vals, double *rhs);

/* ------------------------------------------------------------------ equations
 * FiniteVolumeEquation<T> / CrsEquation held on the device (CrsEquation.cpp,
 * FiniteVolumeEquation.tpp): ncomp * nCells rows in the IndexMap block layout
 * (row = comp * nCells + cell, IndexMap.cpp:31-36), each row its (column,
 * value) pairs in first-touch order, and the stored rhs (A x + rhs = 0).  Every
 * array argument below is a DEVICE pointer (phase_hip_malloc); cell vector
 * fields are interleaved (x0, y0, x1, y1, ...), face fields likewise.  Work is
 * queued on the mesh's stream. */
typedef struct phase_hip_equation phase_hip_equation;

int phase_hip_equation_create(phase_hip_equation **e, phase_hip_mesh *m, int ncomp);
int phase_hip_equation_destroy(phase_hip_equation *e);
int phase_hip_equation_rank(const phase_hip_equation *e);
const char *phase_hip_equation_last_error(const phase_hip_equation *e);

/* Operators: each replaces e's content with a fresh CrsEquation(rank, 5).
 * fv::laplacian, variants as phase_hip_assemble_laplacian (Laplacian.h:7-167;
 * ncomp 2: Laplacian.cpp:5-120 with the SYMMETRY tangential term). */
int phase_hip_eq_laplacian(phase_hip_equation *e, int variant, double gamma, const double *gammaFace,
                           const double *gamma0Face, double theta, const int *bcType, const double *phiFace,
                           const double *phi0Cell, const double *phi0Face);
/* fv::ddt(field, timeStep) -- TimeDerivative.h:37-48 */
int phase_hip_eq_ddt(phase_hip_equation *e, double timeStep, const double *phi0Cell);
/* fv::ddt(Scalar rho, field, timeStep) -- TimeDerivative.h:7-19 (rhoCell NULL)
 * and fv::ddt(rho field, field, timeStep) -- :21-35 (rhoCell, rho0Cell: the
 * density now and at the old time); cellMask (nonzero = in the group, may be
 * NULL) is ddt(field, timeStep, cells) -- :50-62 */
int phase_hip_eq_ddt_rho(phase_hip_equation *e, double timeStep, double rho, const double *rhoCell,
                         const double *rho0Cell, const double *phi0Cell, const int *cellMask);
/* fv::div(u, phi, theta), upwind -- Divergence.h:7-53 (u, u0: face vectors) */
int phase_hip_eq_div(phase_hip_equation *e, const double *uFace, const double *u0Face, double theta,
                     const int *bcType, const double *phiFace, const double *phi0Cell, const double *phi0Face);
/* fv::divc(u, phi, theta), central -- Divergence.h:55-103 */
int phase_hip_eq_divc(phase_hip_equation *e, const double *uFace, const double *u0Face, double theta,
                      const int *bcType, const double *phiFace, const double *phi0Cell, const double *phi0Face);
/* fv::dive(u, phi, theta), explicit (sources only) -- ExplicitDivergence.h:7-51;
 * u0/u1 and phi0/phi1: the two older time levels */
int phase_hip_eq_dive(phase_hip_equation *e, const double *u0Face, const double *u1Face, double theta,
                      const int *bcType, const double *phi0Cell, const double *phi1Cell, const double *phi0Face,
                      const double *phi1Face);
/* fv::div(phiU, gradU, u) -- Divergence.cpp:3-47 (Vector2D equation; gradU a
 * Jacobian per cell, xx xy yx yy) */
int phase_hip_eq_div_grad(phase_hip_equation *e, const double *phiUFace, const double *gradUCell, const int *bcType,
                          const double *uFace);
/* cicsam::div(u, gamma, beta, theta) -- Cicsam.cpp:87-138 (scalar equation;
 * beta from phase_hip_cicsam_weights; gammaFace, gamma0Face: FIXED faces). */
int phase_hip_eq_cicsam_div(phase_hip_equation *e, const double *uFace, const double *beta, double theta,
                            const int *bcType, const double *gammaFace, const double *gamma0Cell,
                            const double *gamma0Face);

/* Algebra: operator+= (sign 1) / -= (sign -1) -- CrsEquation.cpp:185-275;
 * operator*=(Scalar) :287-291; -=(Vector) / ==(Vector) :282-285, 308-311;
 * ==(Scalar) :299-303; implicit under-relaxation a_P /= alpha,
 * rhs -= (1 - alpha) a_P phi (ScalarFiniteVolumeEquation.cpp:45-55); copy. */
int phase_hip_eq_add(phase_hip_equation *e, const phase_hip_equation *o, int sign);
/* operator*(ScalarFiniteVolumeField, FiniteVolumeEquation<T>) -- FiniteVolumeEquation.h:100-102,
 * VectorFiniteVolumeEquation.cpp:164-170: scale(cell, field(cell)) of every
 * cell (CrsEquation::scaleRow, CrsEquation.cpp:153-159; values and rhs);
 * remove(cell) -- ScalarFiniteVolumeEquation.cpp:38-43, VectorFiniteVolumeEquation.cpp:78-91:
 * the masked cells' rows lose their columns and rhs. */
int phase_hip_eq_scale_cells(phase_hip_equation *e, const double *field);
int phase_hip_eq_remove_cells(phase_hip_equation *e, const int *cellMask);
int phase_hip_eq_scale(phase_hip_equation *e, double s);
int phase_hip_eq_sub_vector(phase_hip_equation *e, const double *v);
int phase_hip_eq_eq_scalar(phase_hip_equation *e, double s);
int phase_hip_eq_relax(phase_hip_equation *e, double alpha, const double *phi);
int phase_hip_eq_copy(phase_hip_equation *dst, const phase_hip_equation *src);

/* The equation in the reference's layout (host arrays): rows of a fresh
 * operator have 5 slots (-1 = empty) or more when a row grew; rows rebuilt by
 * += / -= hold exactly their entries.  nslots is queried with rowPtr == NULL. */
int phase_hip_eq_export(phase_hip_equation *e, int *nslots, int *rowPtr, int *colInd, double *vals, double *rhs);

/* FiniteVolumeEquation<T>::solve -- FiniteVolumeEquation.tpp:64-86: the
 * solver's system becomes A = the equation's rows and b = -rhs, device to
 * device (the ordering is rebuilt only when the pattern changed). */
int phase_hip_solver_set_equation(phase_hip_solver *s, const phase_hip_equation *e);

/* x(Index) for every row into a DEVICE array (original numbering), and
 * setGuess from a device array: the field stays in HBM between solves
 * (TrilinosSparseMatrixSolver.cpp:31-38 keeps x across solves). */
int phase_hip_solver_get_x_device(phase_hip_solver *s, double *x);
int phase_hip_solver_set_guess_device(phase_hip_solver *s, const double *x0);

/* Fields (device pointers).  src::div(u) per cell -- Source.cpp:5-21;
 * src::src(phi) = phi V per row, block layout -- Source.cpp:77-95;
 * ScalarGradient::computeFaces + compute(FACE_TO_CELL) -- ScalarGradient.cpp:
 * 34-74; interpolateFaces + setBoundaryFaces -- FiniteVolumeField.tpp:129-182,
 * VectorFiniteVolumeField.cpp:140-160 (FIXED faces keep their value). */
int phase_hip_field_src_div(phase_hip_mesh *m, const double *uFace, double *out);
int phase_hip_field_src_src(phase_hip_mesh *m, int ncomp, const double *phiCell, double *out);
int phase_hip_field_grad(phase_hip_mesh *m, const double *phiCell, const double *phiFace, double *gradFace,
                         double *gradCell);
int phase_hip_field_interpolate_faces(phase_hip_mesh *m, int ncomp, const double *phiCell, const int *bcType,
                                      double *phiFace);
/* VectorFiniteVolumeField::faceToCell -- VectorFiniteVolumeField.cpp:5-26
 * (weights NULL) and its weighted form faceToCell(cellWeight, faceWeight) --
 * :28-51 (both weights given; one without the other is PHASE_HIP_EINVAL). */
int phase_hip_field_face_to_cell(phase_hip_mesh *m, const double *faceVec, const double *cellWeight,
                                 const double *faceWeight, double *cellVec);
/* Volume of fluid, device fields:
 * cicsam::faceInterpolationWeights(u, gamma, gradGamma, timeStep) --
 *   Cicsam.cpp:20-66 (beta per face, 0 on boundary faces);
 * cicsam::computeMomentumFlux(rho1, rho2, u, gamma, beta, rhoU) --
 *   Cicsam.cpp:68-85 (rhoU interleaved per face);
 * the density and viscosity part of FractionalStepMultiphase::updateProperties
 *   -- FractionalStepMultiphase.cpp:175-214 (cells and faces). */
int phase_hip_cicsam_weights(phase_hip_mesh *m, const double *uFace, const double *gammaCell,
                             const double *gradGammaCell, double timeStep, double *beta);
int phase_hip_cicsam_momentum_flux(phase_hip_mesh *m, double rho1, double rho2, const double *uFace,
                                   const double *gammaCell, const double *gammaFace, const double *beta,
                                   double *rhoU);
int phase_hip_vof_properties(phase_hip_mesh *m, double rho1, double rho2, double mu1, double mu2,
                             const double *gammaCell, const double *gammaFace, double *rhoCell, double *rhoFace,
                             double *muCell, double *muFace);
/* Wait for the work queued on the mesh's stream. */
int phase_hip_mesh_synchronize(phase_hip_mesh *m);

/* ------------------------------------------------------------------ ghost-cell immersed boundary
 * GhostCellImmersedBoundary (src/2D/Unstructured/FiniteVolume/ImmersedBoundary/
 * GhostCellImmersedBoundary.cpp:12-90 with its BcStencil, FixedBcStencil and
 * NormalGradientBcStencil) for circles (cx, cy, r per object, in order) or,
 * through phase_hip_gcib_create_shapes, circles and boxes: updateCells classifies the cells (0 fluid, 1 ghost: inside with a
 * face or node neighbour outside, 2 solid) and the ghost stencils are built
 * once (image point, the four cells of the nearest mesh node, their bilinear
 * interpolation; the node needs four cells, as on quad meshes).
 * phase_hip_eq_gcib_bcs is bcs(phi) (:33-70): a fresh scalar equation with
 * the FIXED or NORMAL_GRADIENT stencil rows (source -refValue), unit solid
 * rows (source -refValue for FIXED, 0 for NORMAL_GRADIENT) and empty fluid
 * rows.  phase_hip_eq_gcib_velocity_bcs is velocityBcs(u) (:72-90) for
 * translating objects: FIXED stencils on both components, source minus the
 * object's velocity (2 per object).  A caller assembling over the fluid cells
 * removes the other rows of its fv:: equations (phase_hip_eq_remove_cells
 * with the cell types as mask) and adds these with phase_hip_eq_add.
 * phase_hip_gcib_cell_types copies the types (and owners, host only) to host
 * memory, or the types to device memory when onDevice is nonzero. */
typedef struct phase_hip_gcib phase_hip_gcib;
enum { PHASE_HIP_SHAPE_CIRCLE = 0, PHASE_HIP_SHAPE_BOX = 1 };

int phase_hip_gcib_create(phase_hip_gcib **ib, phase_hip_mesh *m, int nCircles, const double *circles);
/* Bodies of Shape2D::CIRCLE (geometry cx, cy, r, unused) and Shape2D::BOX
 * (lower x, lower y, upper x, upper y; Box.cpp) in order, 4 numbers each. */
int phase_hip_gcib_create_shapes(phase_hip_gcib **ib, phase_hip_mesh *m, int nBodies, const int *kinds,
                                 const double *geometry);
int phase_hip_gcib_destroy(phase_hip_gcib *ib);
const char *phase_hip_gcib_last_error(const phase_hip_gcib *ib);
int phase_hip_gcib_cell_types(const phase_hip_gcib *ib, int *types, int *owner, int onDevice);
int phase_hip_eq_gcib_bcs(phase_hip_equation *e, const phase_hip_gcib *ib, int bcType, double refValue);
int phase_hip_eq_gcib_velocity_bcs(phase_hip_equation *e, const phase_hip_gcib *ib, const double *objectVelocity);
/* velocityBcs(u) for moving bodies, 3 numbers each (vel.x, vel.y, omega): the
 * body velocity at a point is Motion::velocity(pt) = vel + omega (pt -
 * centroid).tangentVec() (Motion.h:31-33), taken at bp for the ghost rows and
 * at the cell centroid for the solid rows. */
int phase_hip_eq_gcib_motion_bcs(phase_hip_equation *e, const phase_hip_gcib *ib, const double *objectMotion);
/* GhostCellImmersedBoundary::applyHydrodynamicForce(rho, mu, u, p, g) --
 * GhostCellImmersedBoundary.cpp:94-146: per body, pressure (bpValue(p) +
 * rho g.bp) and wall shear (mu (grad u(bp) . nw) . nw.tangentVec()) at the
 * boundary points, sorted by angle and integrated edge by edge; uCell
 * (interleaved) and pCell are device fields, force (2 per body) host.
 * Deliberate difference: each body integrates its own boundary points only.
 * The reference keeps one stress list across its body loop (:100-108), so its
 * force on body k also integrates the points of bodies 0..k-1; with one body
 * the two agree. */
int phase_hip_gcib_hydrodynamic_force(phase_hip_gcib *ib, double rho, double mu, const double *uCell,
                                      const double *pCell, double gx, double gy, double *force);

/* ------------------------------------------------------------------ fractional step
 * FractionalStep::solve(timeStep) -- src/2D/Unstructured/Solvers/FractionalStep.cpp:36-117,
 * one time step on the device (all fields in HBM):
 *   uEqn = ddt(u, dt) + div(u, u, 0) == laplacian(mu/rho, u, 0.5) - src(grad p)
 *   u += dt grad p; interpolateFaces(u)
 *   pEqn = laplacian(dt, p) == src::div(u); setBoundaryFaces(p); grad p
 *   u -= dt grad p (cells), u_f -= dt grad_f p (faces)
 * uBcType / pBcType per face (pBcType NULL: zero gradient everywhere);
 * uFace0 / pFace0: initial face values (boundary values of FIXED faces);
 * pinCell >= 0 fixes p at that cell (needed when no pressure face is FIXED),
 * -1 for none.  history (optional, 3 per step): momentum and pressure
 * iterations (summed over the passes), FractionalStep::maxDivergenceError of
 * the corrected faces.  Setup keys u.<k> and p.<k> go to the two solvers;
 * numPressureCorrections = n (1 to 10, default 1) runs the momentum solve,
 * pressure solve and velocity correction n times per step from the same
 * old-time fields, each later pass with the latest face velocities and
 * pressure gradient.  The reference's case files carry this key
 * (Examples/RisingBubble/case/case.info:16) but its fractional-step solvers
 * make one pass, which is the default here. */
typedef struct phase_hip_fstep phase_hip_fstep;

int phase_hip_fstep_create(phase_hip_fstep **s, phase_hip_mesh *m, double rho, double mu, const int *uBcType,
                           const double *uFace0, const int *pBcType, const double *pFace0, int pinCell);
int phase_hip_fstep_destroy(phase_hip_fstep *s);
const char *phase_hip_fstep_last_error(const phase_hip_fstep *s);
int phase_hip_fstep_setup(phase_hip_fstep *s, const char *key, const char *value);
int phase_hip_fstep_advance(phase_hip_fstep *s, double dt, int nSteps, double *history);
int phase_hip_fstep_fields(phase_hip_fstep *s, double *uCell, double *pCell, double *uFace);
/* FractionalStepGCIB::solve -- src/2D/Unstructured/Solvers/FractionalStepGCIB.cpp:14-56:
 * the following steps assemble the fv:: terms on the fluid cells only, add
 * velocityBcs(u) (objectVelocity: 2 per body) and bcs(p) (pBcType FIXED or
 * NORMAL_GRADIENT with pRefValue) of the immersed boundary, use the source
 * src(grad p / rho) and u += dt / rho grad p, and correct and check the
 * divergence on the fluid cells only. */
int phase_hip_fstep_immersed_boundary(phase_hip_fstep *s, const phase_hip_gcib *ib, const double *objectVelocity,
                                      int pBcType, double pRefValue);
/* The same for moving bodies, 3 numbers each (vel.x, vel.y, omega): the
 * velocity rows are velocityBcs(u) with Motion::velocity (Motion.h:31-33), as
 * phase_hip_eq_gcib_motion_bcs builds them. */
int phase_hip_fstep_immersed_boundary_motion(phase_hip_fstep *s, const phase_hip_gcib *ib, const double *objectMotion,
                                             int pBcType, double pRefValue);

/* ------------------------------------------------------------------ multiphase fractional step
 * FractionalStepMultiphase::solve(timeStep) -- src/2D/Unstructured/Solvers/
 * FractionalStepMultiphase.cpp:58-218, one time step on the device; the
 * surface-tension force fst is 0 unless phase_hip_multiphase_surface_tension
 * turned it on:
 *   gamma: ddt(gamma) + cicsam::div(u, gamma, beta, 0.5) == 0 (:76-100)
 *   rho, mu from gamma; sg_f = dot(g, -x_f) grad rho_f; sg = faceToCell (:175-210)
 *   rho ddt(u) + rho dive(u, u, 0.5) == laplacian(mu, u, 0.5) + src(fst + sg - gradP),
 *     u += dt / rho gradP, momentum-interpolated faces (:102-150)
 *   laplacian(dt / rho, p) == src::div(u); gradP = faceToCell(rho, rho) (:152-163)
 *   u -= dt / rho gradP on cells and every face (:165-173)
 * Phase 1 (gamma = 0) has rho1, mu1; phase 2 (gamma = 1) rho2, mu2; (gx, gy)
 * the gravity vector.  Boundary types per face: uBcType (required), pBcType
 * and gammaBcType (NULL: zero gradient everywhere); uFace0, pFace0 boundary
 * values; gammaCell0 the initial volume fraction; pinCell as for
 * phase_hip_fstep_create.  Setup keys gamma.<k>, u.<k>, p.<k> go to the three
 * solvers; numPressureCorrections as for phase_hip_fstep_setup (the passes
 * repeat the momentum, pressure and correction stages, not the volume
 * fraction).  history (optional, 4 per step): gamma, momentum and pressure
 * iterations (summed over the passes), max |src::div(u_f)| after the
 * correction. */
typedef struct phase_hip_multiphase phase_hip_multiphase;

int phase_hip_multiphase_create(phase_hip_multiphase **s, phase_hip_mesh *m, double rho1, double rho2, double mu1,
                                double mu2, double gx, double gy, const int *uBcType, const double *uFace0,
                                const int *pBcType, const double *pFace0, const int *gammaBcType,
                                const double *gammaCell0, int pinCell);
int phase_hip_multiphase_destroy(phase_hip_multiphase *s);
const char *phase_hip_multiphase_last_error(const phase_hip_multiphase *s);
int phase_hip_multiphase_setup(phase_hip_multiphase *s, const char *key, const char *value);
int phase_hip_multiphase_advance(phase_hip_multiphase *s, double dt, int nSteps, double *history);
int phase_hip_multiphase_fields(phase_hip_multiphase *s, double *uCell, double *pCell, double *uFace,
                                double *gammaCell);
/* Turns on the CELESTE surface-tension force of FractionalStepMultiphase
 * (Celeste.cpp:9-109 with CelesteStencil.cpp and SurfaceTensionForce.cpp;
 * input keys Properties.sigma, Solver.smoothingKernelRadius,
 * Solver.kernelType and Boundaries.gamma.<patch>.contactAngle):
 * fst_f = sigma kappa_f gradGamma_f, fst = faceToCell(rho, rho), added to
 * the momentum source and the face-velocity interpolation.  kernelType is a
 * PHASE_HIP_KERNEL_* value; contactAngle per face in radians (NULL: pi/2
 * on every face).  Builds the per-mesh stencils and smoothing kernels and
 * evaluates the force from the current volume fraction, as the reference's
 * initialize() does.  Without this call sigma is 0. */
int phase_hip_multiphase_surface_tension(phase_hip_multiphase *s, double sigma, double kernelRadius,
                                         int kernelType, const double *contactAngle);
/* The model's fields after the last evaluation (host arrays, each may be
 * NULL): smoothed gamma (N), interface normal (2N cells, 2F faces),
 * curvature (N cells, F faces), force (2F faces, 2N cells). */
int phase_hip_multiphase_surface_fields(phase_hip_multiphase *s, double *gammaTilde, double *nCell, double *nFace,
                                        double *kappaCell, double *kappaFace, double *fstFace, double *fstCell);

/* ------------------------------------------------------------------ SIMPLE driver
 * No counterpart module in the reference snapshot (its solvers are projection
 * methods, src/2D/Unstructured/Solvers/FractionalStep.cpp:79-117); built from
 * the README's SIMPLE equations (README.md:27-37) and Patankar's algorithm,
 * with the implicit under-relaxation of ScalarFiniteVolumeEquation.cpp:45-55:
 *   uEqn = rho div(u_f, u) == laplacian(mu, u) - src(grad p), relax(alphaU)
 *   pCorrEqn = laplacian(rho d_f, p') == src::div(u_f)   (Rhie-Chow u_f)
 *   u -= d grad p', u_f -= d_f grad_f p', p += alphaP p'
 * Velocity FIXED on every boundary face (uBcType per face, uBoundary 2F host
 * values); pressure zero-gradient, so p' is pinned to 0 at cell pinCell. */
typedef struct phase_hip_simple phase_hip_simple;

int phase_hip_simple_create(phase_hip_simple **s, phase_hip_mesh *m, double rho, double mu, double alphaU,
                            double alphaP, const int *uBcType, const double *uBoundary, int pinCell);
int phase_hip_simple_destroy(phase_hip_simple *s);
const char *phase_hip_simple_last_error(const phase_hip_simple *s);
/* Inner-solver settings: "u.<key>" / "p.<key>" with phase_hip_solver_setup's keys. */
int phase_hip_simple_setup(phase_hip_simple *s, const char *key, const char *value);
/* nIters outer iterations; history (optional, 4 per iteration): momentum and
 * pressure-correction BiCGStab iterations, max |mass imbalance| before the
 * correction, max |p'|. */
int phase_hip_simple_iterate(phase_hip_simple *s, int nIters, double *history);
/* PISO (Issa 1986) on the same driver, transient: per time step a momentum
 * predictor with fv::ddt(rho, u, dt) and no relaxation, then nCorrectors
 * pressure corrections (Rhie-Chow from the latest u and grad p, p += p').
 * The relaxation factors given at creation are not used.  history as for
 * phase_hip_simple_iterate (pressure iterations summed over the correctors,
 * imbalance and |p'| of the last corrector). */
int phase_hip_simple_piso_advance(phase_hip_simple *s, double dt, int nSteps, int nCorrectors, double *history);
/* Host copies of u (cells, interleaved), p (cells), u (faces); NULL skips. */
int phase_hip_simple_fields(phase_hip_simple *s, double *uCell, double *pCell, double *uFace);

#ifdef __cplusplus
}
#endif
#endif
