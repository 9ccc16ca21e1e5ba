// Sparse solver behind SparseMatrixSolver (src/Math/SparseMatrixSolver.h:11-62):
// right-preconditioned BiCGStab with ILU(0), the method the reference runs
// through Belos + Ifpack2 (src/Math/TrilinosBelosSparseMatrixSolver.cpp:18-86),
// in FP64 or as FP32 inner iterations inside FP64 iterative refinement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "gnn.h"
#include "multilevel.h"
#include "solver.h"
#include "solver_kernels.h"
#include "solver_launch.h"

using namespace phase;

namespace {

thread_local std::string gLastCreateError;

// Scratch of the allocation-free SpMV (phase_hip_solver_spmv_noalloc): the
// operand and the product in solver order, kNoAllocMaxRank rows each, in FP64
// (the FP32 path uses the same bytes).  32 MiB that every GPU the library runs
// on holds until the process ends.  Host code never passes `scratch` itself
// to a kernel or a copy: it uses the device address that
// hipGetSymbolAddress gave for the current device (gScratchPtr), set up when
// a solver is created on that device.
constexpr int kNoAllocMaxRank = 1 << 21;
__device__ __align__(256) double scratch[2 * (size_t)kNoAllocMaxRank];
constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
double *gScratchPtr[kMaxDevices] = {};

// Device address of `scratch` on the current device (which must be `device`),
// looked up once per device.
double *scratchAddress(int device) {
  if (device < 0 || device >= kMaxDevices) fail(PHASE_HIP_EINVAL, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(gScratchMutex);
  if (!gScratchPtr[device]) {
    void *p = nullptr;
    PH_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    gScratchPtr[device] = static_cast<double *>(p);
  }
  return gScratchPtr[device];
}

template <class F>
int guarded(phase_hip_solver *s, F &&f) {
  if (!s) return PHASE_HIP_EINVAL;
  try {
    f();
    return PHASE_HIP_OK;
  } catch (const HipError &e) {
    s->err = e.msg;
    return e.code;
  } catch (const std::exception &e) {
    s->err = e.what();
    return PHASE_HIP_EINVAL;
  }
}

void setDevice(phase_hip_solver *s) { PH_CHECK(hipSetDevice(s->device)); }
template <class T>
void iluApply(phase_hip_solver *s, DBuf<T> &fac, const T *in, T *tmp, T *out, const Scalars *sc) {
  IluDev<T> L = iluDev<T>(s, fac.p);
  launchLU<T, kProNone>(s, L, sc, in, tmp, (T *)nullptr, (const T *)nullptr, (const T *)nullptr, out);
}

// The multigrid preconditioner is in use for this solve.
bool useMl(const phase_hip_solver *s) { return s->precond == 1 && s->ml; }
// Its coarse solves are K cycles (the preconditioner varies from apply to
// apply) or the finest level is smoothed before the correction only (the
// preconditioner is not symmetric): CG runs in its flexible form.
bool flexible(const phase_hip_solver *s) { return useMl(s) && (s->mlKCycle == 1 || !s->mlFinePost); }

// Rest of a preconditioner apply whose fine pre-smoothing (the sweep that
// may also run a Krylov prologue) has written M^-1 rhs into out.
template <class T>
void finishApply(phase_hip_solver *s, const Scalars *sc, const T *rhs, T *out) {
  if (useMl(s)) mlFinishApply<T>(s, sc, rhs, out);
}

// rho' = (r, z), beta ; p = z + beta p  (z in ph)
template <class T>
void cgBeta(phase_hip_solver *s, KrylovVecs<T> &K, Scalars *sc) {
  const int n = s->rank;
  const int gu = gridFor(n, kRedBlock, kRedCap);
  // each direction block re-reads the gu partials: keep the grid near gu
  const int gd = gridFor(n, 4 * kRedBlock, kRedCap);
  if (flexible(s)) {
    hipLaunchKernelGGL((kDotRZ<T, true>), dim3(gu), dim3(kRedBlock), 0, s->stream, n, sc, K.r.p, K.ph.p, K.v.p,
                       s->dPartials.p);
    hipLaunchKernelGGL((kCgDirectionBeta<T, true>), dim3(gd), dim3(kRedBlock), 0, s->stream, n, sc, s->dPartials.p,
                       gu, K.ph.p, K.p.p);
  } else {
    hipLaunchKernelGGL((kDotRZ<T>), dim3(gu), dim3(kRedBlock), 0, s->stream, n, sc, K.r.p, K.ph.p, K.v.p,
                       s->dPartials.p);
    hipLaunchKernelGGL((kCgDirectionBeta<T>), dim3(gd), dim3(kRedBlock), 0, s->stream, n, sc, s->dPartials.p, gu,
                       K.ph.p, K.p.p);
  }
}

// Start of the CG recurrence from r (p is zero): z = M^{-1} r, p = z.
template <class T>
void cgStart(phase_hip_solver *s, const IluDev<T> &L, KrylovVecs<T> &K, Scalars *sc) {
  launchLU<T, kProNone>(s, L, sc, (const T *)K.r.p, K.y.p, (T *)nullptr, (const T *)nullptr, (const T *)nullptr,
                        K.ph.p);
  finishApply<T>(s, sc, K.r.p, K.ph.p);
  cgBeta<T>(s, K, sc);
}

// One right-preconditioned BiCGStab iteration, all scalars on the device.
template <class T, class TX, class TV>
void iterate(phase_hip_solver *s, KrylovVecs<T> &K, DBuf<T> &fac, const TV *sell, TX *x, Scalars *sc) {
  const int n = s->rank;
  IluDev<T> L = iluDev<T>(s, fac.p);
  // p = r + beta (p - omega v) ; ph = M^{-1} p
  launchLU<T, kProP>(s, L, sc, (const T *)nullptr, K.y.p, K.p.p, K.r.p, K.v.p, K.ph.p);
  finishApply<T>(s, sc, K.p.p, K.ph.p);
  // v = A ph ; alpha
  const int gs = spmvGrid(s);
  hipLaunchKernelGGL((kSpmvDotV<T, TV>), dim3(gs), dim3(kRedBlock), 0, s->stream, n, s->ord.nSlices, s->dSliceOff.p,
                     s->dSellCol.p, sell, K.ph.p, K.v.p, K.r0.p, sc, s->dPartials.p);
  hipLaunchKernelGGL(kFinishV, dim3(1), dim3(kFinBlock), 0, s->stream, sc, s->dPartials.p, gs);
  // s = r - alpha v ; sh = M^{-1} s
  launchLU<T, kProS>(s, L, sc, (const T *)nullptr, K.y.p, K.s.p, K.r.p, K.v.p, K.sh.p);
  finishApply<T>(s, sc, K.s.p, K.sh.p);
  // t = A sh ; omega
  hipLaunchKernelGGL((kSpmvDotT<T, TV>), dim3(gs), dim3(kRedBlock), 0, s->stream, n, s->ord.nSlices, s->dSliceOff.p,
                     s->dSellCol.p, sell, K.sh.p, K.t.p, K.s.p, sc, s->dPartials.p);
  hipLaunchKernelGGL(kFinishT, dim3(1), dim3(kFinBlock), 0, s->stream, sc, s->dPartials.p, gs);
  // x, r, rho, beta, convergence
  const int gu = gridFor(n, kRedBlock, kRedCap);
  hipLaunchKernelGGL((kUpdate<T, TX>), dim3(gu), dim3(kRedBlock), 0, s->stream, n, sc, x, K.ph.p, K.sh.p, K.r.p,
                     K.s.p, K.t.p, K.r0.p, s->dPartials.p);
  hipLaunchKernelGGL(kFinishUpdate, dim3(1), dim3(kRedBlock), 0, s->stream, sc, s->dPartials.p, gu);
}

// One preconditioned CG iteration: q = A p, the stopping test and alpha; x and
// r moved inside the sweeps that form z = M^{-1} r; beta; p.
template <class T, class TX, class TV>
void iterateCg(phase_hip_solver *s, KrylovVecs<T> &K, DBuf<T> &fac, const TV *sell, TX *x, Scalars *sc) {
  static_assert(std::is_same<TX, double>::value, "the CG prologue updates an FP64 solution");
  const int n = s->rank;
  IluDev<T> L = iluDev<T>(s, fac.p);
  const int gs = spmvGrid(s);
  hipLaunchKernelGGL((kSpmvDotCg<T, TV>), dim3(gs), dim3(kRedBlock), 0, s->stream, n, s->ord.nSlices,
                     s->dSliceOff.p, s->dSellCol.p, sell, K.p.p, K.v.p, K.r.p, sc, s->dPartials.p);
  hipLaunchKernelGGL(kFinishCgQ, dim3(1), dim3(kFinBlock), 0, s->stream, sc, s->dPartials.p, gs);
  launchLU<T, kProCgR>(s, L, sc, (const T *)K.p.p, K.y.p, K.r.p, (const T *)nullptr, (const T *)K.v.p, K.ph.p, x);
  finishApply<T>(s, sc, K.r.p, K.ph.p);
  cgBeta<T>(s, K, sc);
}

double normB2(phase_hip_solver *s) {
  const int n = s->rank;
  const int g = gridFor(n, kRedBlock, kRedCap);
  hipLaunchKernelGGL((kNorm2<double>), dim3(g), dim3(kRedBlock), 0, s->stream, n, s->dB.p, s->dPartials.p);
  hipLaunchKernelGGL(kFinishNorm2, dim3(1), dim3(kFinBlock), 0, s->stream, s->dPartials.p, g, s->dMisc.p);
  double b2 = 0.;
  PH_CHECK(hipMemcpyAsync(&b2, s->dMisc.p, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  PH_CHECK(hipStreamSynchronize(s->stream));
  return b2;
}

// Run Krylov chunks until the device says done; returns iterations taken.
// onChunk(iterationsSoFar) runs on the host after every chunk and returns
// false to stop the recurrence.  Iterations queued after the one that meets
// the threshold still launch (each returns at once), so once a rate is known
// the next chunk is the number of iterations the last chunk's rate of
// decrease needs to reach the threshold (at most `chunk`): a guess one short
// costs a host round trip, one long an iteration of empty launches.
template <class T, class TX, class TV, class OnChunk>
int runKrylov(phase_hip_solver *s, KrylovVecs<T> &K, DBuf<T> &fac, const TV *sell, TX *x, int budget,
              OnChunk &&onChunk) {
  Scalars *sc = s->dSc.p;  // slot 0
  int it = 0, chunk = s->chunk, itLast = 0;
  double rrLast = -1.;
  while (true) {
    for (int c = 0; c < chunk; ++c) {
      if (s->cg) iterateCg<T, TX, TV>(s, K, fac, sell, x, sc);
      else iterate<T, TX, TV>(s, K, fac, sell, x, sc);
    }
    PH_CHECK(hipGetLastError());
    pullScalars(s);
    const Scalars &h = s->hSc[0];
    it = h.iters;
    if (h.done != kDoneRunning || it >= budget) break;
    if (!onChunk(it)) break;
    chunk = s->chunk;
    if (rrLast > 0. && h.rr > h.tol2 && h.rr < rrLast && it > itLast) {
      const double need = std::log(h.tol2 / h.rr) / (std::log(h.rr / rrLast) / (it - itLast));
      if (need < chunk) chunk = std::max(1, (int)std::ceil(need));
    }
    rrLast = h.rr;
    itLast = it;
  }
  checkDeviceErr(s);
  return it;
}

bool capturing(phase_hip_solver *s) { return s->gnn && s->gnn->captureEvery > 0; }

// FP64 BiCGStab + ILU(0) from the current x.  `total` iterations were already
// spent by the caller (a mixed solve falling back); the loop restarts the
// recurrence from the FP64 true residual after a breakdown.
void fp64Loop(phase_hip_solver *s, double b2, int total) {
  const int n = s->rank;
  ensureFactor64(s);
  KrylovVecs<double> &K = s->k64;
  K.alloc(n);
  const double tol2 = s->tol * s->tol * b2;
  Scalars *sc = s->dSc.p;
  while (true) {
    // r = b - A x (true residual), r0 = r
    hipLaunchKernelGGL((kResidual<double>), dim3(spmvGrid(s)), dim3(kRedBlock), 0, s->stream, n,
                       s->ord.nSlices, s->dSliceOff.p, s->dSellCol.p, s->dSell64.p, s->dX.p, s->dB.p, K.r.p,
                       K.r0.p, s->dPartials.p);
    hipLaunchKernelGGL(kFinishResidual, dim3(1), dim3(kFinBlock), 0, s->stream, sc, s->dPartials.p, spmvGrid(s));
    hipLaunchKernelGGL((kBicgInit<double>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n, sc,
                       K.p.p, K.v.p, (double *)nullptr, tol2, s->maxIters - total, 0.0);
    if (s->cg) cgStart<double>(s, iluDev<double>(s, s->dFac64.p), K, sc);
    pullScalars(s);
    s->error = std::sqrt(s->hSc[0].rr / b2);
    ++s->outerPasses;
    if (s->hSc[0].done == kDoneConverged || total >= s->maxIters || !std::isfinite(s->error)) break;
    if (capturing(s) && total == 0) {
      gnnCaptureIter(s, K.r.p, s->dX.p, 0);
      s->lastCaptureIter = 0;
    }
    const int base = total;
    int it = runKrylov<double, double, double>(s, K, s->dFac64, s->dSell64.p, s->dX.p, s->maxIters - total,
                                               [&](int sofar) {
      if (capturing(s) && base + sofar - s->lastCaptureIter >= s->gnn->captureEvery) {
        gnnCaptureIter(s, K.r.p, s->dX.p, base + sofar);
        s->lastCaptureIter = base + sofar;
      }
      return true;
    });
    total += it;
    if (s->hSc[0].done == kDoneBreakdown && it == 0) break;
  }
  s->innerIters = total;
  s->nIters = total;
}

void solveFp64(phase_hip_solver *s) {
  const int n = s->rank;
  const double b2 = normB2(s);
  s->innerIters = 0;
  s->outerPasses = 0;
  if (b2 == 0.) {
    PH_CHECK(hipMemsetAsync(s->dX.p, 0, n * sizeof(double), s->stream));
    s->error = 0.;
    return;
  }
  fp64Loop(s, b2, 0);
  if (capturing(s)) gnnCaptureFinish(s, s->dX.p, s->error <= s->tol);
}

// Learned step of a refinement pass: e = network(r); the step is the
// minimal-residual multiple alpha e of it, kept only when it reduces ||r|| by
// at least the fallback ratio -- otherwise the rest of the solve runs without
// the network (plain ILU(0) refinement).
bool learnedStep(phase_hip_solver *s, double r2) {
  const int n = s->rank;
  s->dE64.alloc(n);
  s->dQ64.alloc(n);
  Scalars *misc = s->dSc.p + 2, *outer = s->dSc.p + 1;
  gnnApply(s, s->dR64.p, s->dE64.p);
  hipLaunchKernelGGL((kSpmvDot2<double>), dim3(spmvGrid(s)), dim3(kRedBlock), 0, s->stream, n, s->ord.nSlices,
                     s->dSliceOff.p, s->dSellCol.p, s->dSell64.p, s->dE64.p, s->dQ64.p, s->dR64.p, misc,
                     s->dPartials.p, s->dCounter.p);
  pullScalars(s);
  const double rq = s->hSc[2].ts, qq = s->hSc[2].tt;
  if (!(qq > 0.) || !std::isfinite(rq) || !std::isfinite(qq)) return false;
  const double newr2 = r2 - rq * rq / qq;
  const double ratio = std::sqrt(std::max(newr2, 0.) / r2);
  if (!(ratio <= s->fallbackRatio)) return false;
  hipLaunchKernelGGL((kMinResUpdate<double>), dim3(gridFor(n, kRedBlock, kRedCap)), dim3(kRedBlock), 0, s->stream, n,
                     misc, outer, s->dX.p, s->dE64.p, s->dR64.p, s->dQ64.p, s->dPartials.p, s->dCounter.p);
  PH_CHECK(hipGetLastError());
  return true;
}

// Mixed precision: BiCGStab iterates with FP32 Krylov vectors and an FP32
// ILU(0) factor, while the matrix coefficients stay FP64 (each SpMV row
// accumulates in FP64 and rounds once), all dot products reduce in FP64 and
// the solution accumulates in FP64.  Whenever the recursive residual has
// dropped by `innerTolerance` since the last refinement point, r is replaced by
// the FP64 true residual b - A x (a reliable update: FP64 iterative refinement
// that keeps the Krylov recurrence).  Convergence is judged on the FP64 true
// residual only.  The iterate with the smallest true residual is kept; if the
// true residual grows a hundredfold past it, stops being finite, or fails to
// halve over kStallPasses refinement points in a row, or the FP32 recursive
// residual makes no new low for twice the longest stretch between refinement
// points so far (at least kStallMinIters iterations; FP32 stagnation), the
// solve falls back to FP64 from that iterate.
void solveMixed(phase_hip_solver *s) {
  const int n = s->rank;
  ensureSingle(s);
  KrylovVecs<float> &K = s->k32;
  K.alloc(n);
  s->dR64.alloc(n);
  s->dXBest.alloc(n);
  const double b2 = normB2(s);
  s->innerIters = 0;
  s->outerPasses = 0;
  if (b2 == 0.) {
    PH_CHECK(hipMemsetAsync(s->dX.p, 0, n * sizeof(double), s->stream));
    s->error = 0.;
    return;
  }
  bool useNet = s->mode == PHASE_HIP_MODE_LEARNED && s->gnn && s->gnn->trained;
  if (useNet) gnnPrepare(s);
  Scalars *inner = s->dSc.p, *outer = s->dSc.p + 1;
  const double tol2 = s->tol * s->tol * b2;
  const double delta2 = s->innerTol * s->innerTol;
  double bestRR = HUGE_VAL;
  bool diverged = false;
  constexpr int kStallPasses = 4, kStallMinIters = 800;
  int stall = 0, longestGap = 0;
  // true residual into dR64 / outer->rr; keeps the best iterate
  auto trueResidual = [&]() {
    hipLaunchKernelGGL((kResidual<double>), dim3(spmvGrid(s)), dim3(kRedBlock), 0, s->stream, n,
                       s->ord.nSlices, s->dSliceOff.p, s->dSellCol.p, s->dSell64.p, s->dX.p, s->dB.p,
                       s->dR64.p, (double *)nullptr, s->dPartials.p);
    hipLaunchKernelGGL(kFinishResidual, dim3(1), dim3(kFinBlock), 0, s->stream, outer, s->dPartials.p,
                       spmvGrid(s));
    pullScalars(s);
    const double rr = s->hSc[1].rr;
    if (!(rr <= 1e4 * bestRR)) {  // NaN included
      diverged = true;
      return rr;
    }
    stall = rr < 0.25 * bestRR ? 0 : stall + 1;  // ||r|| at least halved, or not
    if (stall >= kStallPasses) {
      diverged = true;
      return rr;
    }
    if (rr < bestRR) {
      bestRR = rr;
      PH_CHECK(hipMemcpyAsync(s->dXBest.p, s->dX.p, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    }
    return rr;
  };
  int total = 0;
  for (int pass = 0; pass <= s->maxOuter; ++pass) {
    const double rr = trueResidual();
    s->error = std::sqrt(rr / b2);
    if (diverged || rr <= tol2 || total >= s->maxIters || pass == s->maxOuter) break;
    ++s->outerPasses;
    if (useNet) {
      if (learnedStep(s, rr)) {
        ++s->learnedSteps;
      } else {
        useNet = false;
        ++s->fallbacks;
      }
    }
    // restart the FP32 recurrence from the FP64 residual
    hipLaunchKernelGGL((kConvert2<double, float>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                       s->dR64.p, K.r.p, K.r0.p);
    hipLaunchKernelGGL((kNorm2<float>), dim3(gridFor(n, kRedBlock, kRedCap)), dim3(kRedBlock), 0, s->stream,
                       n, K.r.p, s->dPartials.p);
    hipLaunchKernelGGL(kFinishNorm2, dim3(1), dim3(kFinBlock), 0, s->stream, s->dPartials.p,
                       gridFor(n, kRedBlock, kRedCap), &inner->rr);
    hipLaunchKernelGGL((kBicgInit<float>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n, inner,
                       K.p.p, K.v.p, (float *)nullptr, tol2, s->maxIters - total, 0.0);
    if (s->cg) cgStart<float>(s, iluDev<float>(s, s->dFac32.p), K, inner);
    pullScalars(s);
    double lastRU = s->hSc[0].rr, minRU = lastRU;
    int lastUpdate = 0, minAt = 0;
    int it = runKrylov<float, double, double>(s, K, s->dFac32, s->dSell64.p, s->dX.p, s->maxIters - total,
                                              [&](int sofar) {
      if (!(s->hSc[0].rr < delta2 * lastRU)) {
        // the FP32 recurrence makes no new low for a long stretch before the
        // next refinement point: stagnation, as when the true residual stops halving
        if (s->hSc[0].rr < minRU) {
          minRU = s->hSc[0].rr;
          minAt = sofar;
        } else if (sofar - minAt > std::max(kStallMinIters, 2 * longestGap)) {
          diverged = true;
          return false;
        }
        return true;
      }
      longestGap = std::max(longestGap, sofar - lastUpdate);
      lastUpdate = minAt = sofar;
      // reliable update: r = b - A x in FP64, handed to the FP32 recurrence
      trueResidual();
      if (diverged) return false;
      hipLaunchKernelGGL((kReplaceResidual<float>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream,
                         n, s->dR64.p, K.r.p, inner, outer);
      lastRU = minRU = s->hSc[1].rr;
      ++s->outerPasses;
      return true;
    });
    total += it;
    if (diverged || it == 0) break;
  }
  if (diverged || !(s->error <= s->tol)) {
    // not converged in FP32: restore the best iterate and, while the budget
    // allows, finish in FP64 from it (divergence, a breakdown right after a
    // restart, or the refinement-pass limit all end up here)
    if (bestRR < HUGE_VAL && (diverged || bestRR < s->error * s->error * b2)) {
      PH_CHECK(hipMemcpyAsync(s->dX.p, s->dXBest.p, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
      s->error = std::sqrt(bestRR / b2);
    }
    if (total < s->maxIters) {
      ++s->fallbacks;
      fp64Loop(s, b2, total);
      return;
    }
  }
  s->innerIters = total;
  s->nIters = total;
}

void allocCommon(phase_hip_solver *s) {
  const int n = s->rank;
  s->dB.alloc(n);
  s->dX.alloc(n);
  s->dErr.alloc(1);
  s->dErr.zero(s->stream);
  const int maxGrid = std::max(spmvGrid(s), kRedCap);
  s->dPartials.alloc(4 * (size_t)maxGrid);
  s->dCounter.alloc(1);
  s->dCounter.zero(s->stream);
  s->dSc.alloc(4);  // [3]: idle slot, done == 0, for sweeps outside a Krylov run
  s->dSc.zero(s->stream);
  s->dMisc.alloc(4);
  if (!s->hSc) PH_CHECK(hipHostMalloc((void **)&s->hSc, 3 * sizeof(Scalars), hipHostMallocDefault));
}

}  // namespace

namespace phase {

void solverSetPattern(phase_hip_solver *s, int n, const int *rowPtr, const int *colInd, const double *cellXY) {
  setDevice(s);
  if (s->rank != 0 && n != s->rank) fail(PHASE_HIP_EINVAL, "pattern size differs from setRank");
  buildOrdering(s->ord, n, rowPtr, colInd, s->blockRows, s->geometric ? cellXY : nullptr, s->blockJacobi != 0,
                s->iluLimits != 0);
  mlFree(s);
  if (cellXY && s->geometric) s->cellXY.assign(cellXY, cellXY + 2 * (size_t)n);
  else s->cellXY.clear();
  s->rank = n;
  const Ordering &o = s->ord;
  hipStream_t st = s->stream;
  s->dBlocks.upload(o.blocks, st);
  s->dRows.upload(o.rows, st);
  s->dLanes.upload(o.lanes, st);
  s->dLCol.upload(o.lCol.empty() ? std::vector<int>{-1} : o.lCol, st);
  s->dUCol.upload(o.uCol.empty() ? std::vector<int>{-1} : o.uCol, st);
  s->dUpdPtr.upload(o.updPtr, st);
  s->dUpdSrc.upload(o.updSrc.empty() ? std::vector<int>{0} : o.updSrc, st);
  s->dUpdDst.upload(o.updDst.empty() ? std::vector<int>{0} : o.updDst, st);
  s->dSliceOff.upload(o.sliceOff, st);
  s->dSellCol.upload(o.sellCol, st);
  s->dSellSrc.upload(o.sellSrc, st);
  s->dFacSrc.upload(o.facSrc, st);
  s->dPerm.upload(o.perm, st);
  s->dSell64.alloc(o.sellCol.size());
  s->eqStride = 0;  // the pattern no longer comes from an equation (set_equation re-marks it)
  s->patRowPtr.assign(rowPtr, rowPtr + n + 1);
  s->patColInd.assign(colInd, colInd + rowPtr[n]);
  s->dVals.alloc(std::max<size_t>(1, (size_t)rowPtr[n]));
  s->havePattern = true;
  s->haveValues = false;
  s->fac64Valid = s->fac32Valid = s->sell32Valid = false;
  allocCommon(s);
  PH_CHECK(hipStreamSynchronize(st));
}

void solverEnsureSingle(phase_hip_solver *s) { ensureSingle(s); }

}  // namespace phase

namespace {
// Plain copy, 16 bytes per lane per step: one read and one write of every byte.
__global__ void kCopy16(size_t n16, const uint4 *__restrict__ src, uint4 *__restrict__ dst) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

__global__ void kCmpInt(size_t n, const int *__restrict__ a, const int *__restrict__ b, int *flag) {
  bool diff = false;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    diff |= a[i] != b[i];
  if (diff) atomicOr(flag, 1);
}
// flag = 1 when a column of a new row (stride ns) is missing from the stored
// row (stride os)
__global__ void kSubsetCols(int rank, int ns, const int *__restrict__ nc, int os, const int *__restrict__ oc,
                            int *flag) {
  bool missing = false;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rank; r += gridDim.x * blockDim.x)
    for (int j = 0; j < ns; ++j) {
      const int c = nc[(size_t)r * ns + j];
      if (c < 0) continue;
      bool found = false;
      for (int q = 0; q < os && !found; ++q) found = oc[(size_t)r * os + q] == c;
      missing |= !found;
    }
  if (missing) atomicOr(flag, 1);
}

// the new rows' values in the stored layout, 0 where a new row has no entry
__global__ void kRemapVals(int rank, int ns, const int *__restrict__ nc, const double *__restrict__ nv, int os,
                           const int *__restrict__ oc, double *__restrict__ ov) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rank; r += gridDim.x * blockDim.x)
    for (int q = 0; q < os; ++q) {
      const int c = oc[(size_t)r * os + q];
      double v = 0.;
      if (c >= 0)
        for (int j = 0; j < ns; ++j)
          if (nc[(size_t)r * ns + j] == c) {
            v = nv[(size_t)r * ns + j];
            break;
          }
      ov[(size_t)r * os + q] = v;
    }
}
}  // namespace

namespace phase {

void solverSetFromRows(phase_hip_solver *s, int rank, int stride, const int *dCols, const double *dVals,
                       const double *dRhs, hipStream_t producer, const double *dCellXY, const int *adjPtr,
                       const int *adjNb) {
  setDevice(s);
  if (rank <= 0) fail(PHASE_HIP_EINVAL, "empty equation");
  PH_CHECK(hipStreamSynchronize(producer));
  const size_t slots = (size_t)rank * stride;
  bool same = s->havePattern && !s->binding && s->rank == rank && s->eqStride == stride && s->dEqCols.n == slots;
  if (same) {
    // the pattern follows the values (CrsEquation drops zeros): compare on the device
    s->dEqFlag.alloc(1);
    s->dEqFlag.zero(s->stream);
    hipLaunchKernelGGL(kCmpInt, dim3(gridFor((long)slots, 256, kElemCap)), dim3(256), 0, s->stream, slots, dCols,
                       s->dEqCols.p, s->dEqFlag.p);
    int flag = 0;
    PH_CHECK(hipMemcpyAsync(&flag, s->dEqFlag.p, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    PH_CHECK(hipStreamSynchronize(s->stream));
    same = flag == 0;
  }
  // union patterns: a pattern inside the stored one keeps the ordering and
  // factor structure, with explicit zeros for the entries it lacks
  const bool canUnion = s->unionPattern && s->havePattern && !s->binding && s->rank == rank && s->eqStride > 0;
  bool remap = false;
  if (!same && canUnion) {
    s->dEqFlag.alloc(1);
    s->dEqFlag.zero(s->stream);
    hipLaunchKernelGGL(kSubsetCols, dim3(gridFor(rank, 256, kElemCap)), dim3(256), 0, s->stream, rank, stride, dCols,
                       s->eqStride, s->dEqCols.p, s->dEqFlag.p);
    int flag = 0;
    PH_CHECK(hipMemcpyAsync(&flag, s->dEqFlag.p, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    PH_CHECK(hipStreamSynchronize(s->stream));
    remap = flag == 0;
  }
  if (!same && !remap) {
    std::vector<int> cols(slots), rowPtr(rank + 1);
    PH_CHECK(hipMemcpy(cols.data(), dCols, slots * sizeof(int), hipMemcpyDeviceToHost));
    int useStride = stride;
    if (s->unionPattern && (canUnion || adjPtr)) {
      // the new entries, then the stored ones and the structural neighbours
      // they lack: later patterns inside this union only remap values
      std::vector<int> old;
      if (canUnion) {
        old.resize((size_t)rank * s->eqStride);
        PH_CHECK(hipMemcpy(old.data(), s->dEqCols.p, old.size() * sizeof(int), hipMemcpyDeviceToHost));
      }
      std::vector<int> width(rank), merged;
      merged.reserve((size_t)rank * (stride + 8));
      std::vector<int> start(rank + 1, 0);
      auto add = [&](size_t from, int c) {
        if (c < 0) return;
        for (size_t q = from; q < merged.size(); ++q)
          if (merged[q] == c) return;
        merged.push_back(c);
      };
      int wmax = 1;
      for (int r = 0; r < rank; ++r) {
        const size_t from = merged.size();
        for (int j = 0; j < stride; ++j) add(from, cols[(size_t)r * stride + j]);
        if (canUnion)
          for (int q = 0; q < s->eqStride; ++q) add(from, old[(size_t)r * s->eqStride + q]);
        if (adjPtr) {
          add(from, r);
          for (int q = adjPtr[r]; q < adjPtr[r + 1]; ++q) add(from, adjNb[q]);
        }
        start[r + 1] = (int)merged.size();
        wmax = std::max(wmax, start[r + 1] - start[r]);
      }
      cols.assign((size_t)rank * wmax, -1);
      for (int r = 0; r < rank; ++r)
        std::copy(merged.begin() + start[r], merged.begin() + start[r + 1], cols.begin() + (size_t)r * wmax);
      useStride = wmax;
      remap = true;
    }
    std::vector<double> xy;
    if (dCellXY && s->geometric) {
      xy.resize(2 * (size_t)rank);
      PH_CHECK(hipMemcpy(xy.data(), dCellXY, xy.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    for (int i = 0; i <= rank; ++i) rowPtr[i] = i * useStride;
    // the last solution survives a new ordering of the same rows (warm start)
    DBuf<double> keep;
    const bool warm = s->havePattern && s->rank == rank && (s->haveSolution || s->haveGuess);
    if (warm) {
      keep.alloc(rank);
      hipLaunchKernelGGL((kPermuteOut<double>), dim3(gridFor(rank, 256, kElemCap)), dim3(256), 0, s->stream, rank,
                         s->dPerm.p, s->dX.p, keep.p);
      PH_CHECK(hipStreamSynchronize(s->stream));
    }
    solverFreeBinding(s);
    s->rank = 0;
    solverSetPattern(s, rank, rowPtr.data(), cols.data(), xy.empty() ? nullptr : xy.data());
    if (warm) {
      hipLaunchKernelGGL((kPermuteIn<double, double>), dim3(gridFor(rank, 256, kElemCap)), dim3(256), 0, s->stream,
                         rank, s->dPerm.p, keep.p, s->dX.p, 1.0);
      PH_CHECK(hipStreamSynchronize(s->stream));  // before `keep` is freed
    }
    s->dEqCols.alloc((size_t)rank * useStride);  // after solverSetPattern, which clears eqStride
    if (remap)
      PH_CHECK(hipMemcpy(s->dEqCols.p, cols.data(), cols.size() * sizeof(int), hipMemcpyHostToDevice));
    else
      PH_CHECK(hipMemcpyAsync(s->dEqCols.p, dCols, slots * sizeof(int), hipMemcpyDeviceToDevice, s->stream));
    s->eqStride = useStride;
  }
  if (remap)
    hipLaunchKernelGGL(kRemapVals, dim3(gridFor(rank, 256, kElemCap)), dim3(256), 0, s->stream, rank, stride, dCols,
                       dVals, s->eqStride, s->dEqCols.p, s->dVals.p);
  else
    PH_CHECK(hipMemcpyAsync(s->dVals.p, dVals, slots * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
  solverValuesChanged(s);
  hipLaunchKernelGGL((kPermuteIn<double, double>), dim3(gridFor(rank, 256, kElemCap)), dim3(256), 0, s->stream,
                     rank, s->dPerm.p, dRhs, s->dB.p, -1.0);
  PH_CHECK(hipGetLastError());
  s->haveRhs = true;
}

template <class T>
void iluApplyDev(phase_hip_solver *s, T *fac, const T *in, T *tmp, T *out) {
  IluDev<T> L = iluDev<T>(s, fac);
  launchLU<T, kProNone>(s, L, idleScalars(s), in, tmp, (T *)nullptr, (const T *)nullptr, (const T *)nullptr,
                        out);
}
template void iluApplyDev<float>(phase_hip_solver *, float *, const float *, float *, float *);
template void iluApplyDev<double>(phase_hip_solver *, double *, const double *, double *, double *);

void solverValuesChanged(phase_hip_solver *s) {
  const int m = (int)s->dSell64.n;
  hipLaunchKernelGGL((kGather<double>), dim3(gridFor(m, 256, kElemCap)), dim3(256), 0, s->stream, m,
                     s->dSellSrc.p, s->dVals.p, s->dSell64.p);
  PH_CHECK(hipGetLastError());
  s->haveValues = true;
  s->fac64Valid = s->fac32Valid = s->sell32Valid = false;
  mlValuesChanged(s);
}

}  // namespace phase

extern "C" {

const char *phase_hip_last_error(const phase_hip_solver *s) {
  return s ? s->err.c_str() : gLastCreateError.c_str();
}

int phase_hip_device_count(int *count) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  if (count) *count = c;
  return PHASE_HIP_OK;
}

int phase_hip_solver_create(phase_hip_solver **out, int device) {
  if (!out) return PHASE_HIP_EINVAL;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    gLastCreateError = "no HIP device available";
    return PHASE_HIP_EDEVICE;
  }
  if (device < 0 || device >= count) {
    gLastCreateError = "device index out of range";
    return PHASE_HIP_EINVAL;
  }
  phase_hip_solver *s = new phase_hip_solver();
  s->device = device;
  int rc = guarded(s, [&] {
    setDevice(s);
    scratchAddress(device);  // the allocation-free entry points' scratch on this device
    PH_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    PH_CHECK(hipEventCreate(&s->ev0));
    PH_CHECK(hipEventCreate(&s->ev1));
  });
  if (rc != PHASE_HIP_OK) {
    gLastCreateError = s->err;
    delete s;
    return rc;
  }
  *out = s;
  return PHASE_HIP_OK;
}

int phase_hip_solver_destroy(phase_hip_solver *s) {
  if (!s) return PHASE_HIP_OK;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  solverFreeBinding(s);
  gnnFree(s);
  mlFree(s);
  s->dBlocks.release();
  if (s->hSc) (void)hipHostFree(s->hSc);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  // device buffers are released by their destructors before the stream goes
  hipStream_t st = s->stream;
  delete s;
  if (st) (void)hipStreamDestroy(st);
  return PHASE_HIP_OK;
}

int phase_hip_solver_setup(phase_hip_solver *s, const char *key, const char *value) {
  return guarded(s, [&] {
    if (!key || !value) fail(PHASE_HIP_EINVAL, "null key or value");
    std::string k(key), v(value);
    std::string lv = v;
    std::transform(lv.begin(), lv.end(), lv.begin(), ::tolower);
    if (k == "tolerance") s->tol = std::stod(v);
    else if (k == "maxIters") s->maxIters = std::stoi(v);
    else if (k == "mode") {
      if (lv == "fp64" || lv == "0") s->mode = PHASE_HIP_MODE_FP64;
      else if (lv == "mixed" || lv == "1") s->mode = PHASE_HIP_MODE_MIXED;
      else if (lv == "learned" || lv == "2") s->mode = PHASE_HIP_MODE_LEARNED;
      else fail(PHASE_HIP_EINVAL, "unknown mode '" + v + "'");
    } else if (k == "innerTolerance") s->innerTol = std::stod(v);
    else if (k == "blockRows") {
      int br = std::stoi(v);
      if (br < 1 || br > 64) fail(PHASE_HIP_EINVAL, "blockRows must be in [1,64]");
      if (br != s->blockRows && s->havePattern) s->havePattern = false;  // rebuild on next set
      s->blockRows = br;
    } else if (k == "fallbackRatio") s->fallbackRatio = std::stod(v);
    else if (k == "maxOuter") s->maxOuter = std::stoi(v);
    else if (k == "chunk") s->chunk = std::max(1, std::stoi(v));
    else if (k == "solver") {
      if (lv == "bicgstab") s->cg = 0, s->gmres = 0;
      else if (lv == "cg") s->cg = 1, s->gmres = 0;
      else if (lv == "gmres") s->cg = 0, s->gmres = 1;
      else fail(PHASE_HIP_EINVAL, "unknown solver '" + v + "' (BICGSTAB, CG or GMRES)");
    } else if (k == "gmresRestart") {
      const int r = std::stoi(v);
      if (r < 1 || r > kGmMaxRestart) fail(PHASE_HIP_EINVAL, "gmresRestart must be in [1,64]");
      s->gmresRestart = r;
    } else if (k == "ordering") {
      const int g2 = lv == "geometric" ? 1 : lv == "graph" ? 0 : -1;
      if (g2 < 0) fail(PHASE_HIP_EINVAL, "unknown ordering '" + v + "' (geometric or graph)");
      if (g2 != s->geometric && s->havePattern) s->havePattern = false;  // rebuild on next set
      s->geometric = g2;
    } else if (k == "preconditioner") {
      int pc = lv == "ilu" || lv == "riluk" || lv == "ilu0" ? 0 : lv == "multigrid" || lv == "amg" ? 1 : -1;
      if (pc < 0) fail(PHASE_HIP_EINVAL, "unknown preconditioner '" + v + "' (ILU or MULTIGRID)");
      s->precond = pc;
    } else if (k == "multigridCycle") {
      int kc = lv == "k" ? 1 : lv == "v" ? 0 : lv == "auto" ? -1 : -2;
      if (kc < -1) fail(PHASE_HIP_EINVAL, "unknown multigrid cycle '" + v + "' (K, V or auto)");
      s->mlKCycleSetting = kc;
    } else if (k == "multigridKLevels") {
      const int r = std::stoi(v);
      if (r < 0) fail(PHASE_HIP_EINVAL, "multigridKLevels must be >= 0");
      s->mlKLevels = r;
    } else if (k == "multigridBlockJacobiFrom") {
      const int r = std::stoi(v);
      if (r < 0) fail(PHASE_HIP_EINVAL, "multigridBlockJacobiFrom must be >= 0");
      if (r != s->mlBlockJacobiFrom) mlFree(s);
      s->mlBlockJacobiFrom = r;
    } else if (k == "multigridJacobiFrom") {
      const int r = lv == "auto" ? -1 : std::stoi(v);
      if (r < -1) fail(PHASE_HIP_EINVAL, "multigridJacobiFrom must be AUTO or >= 0");
      s->mlJacobiFrom = r;  // the hierarchy is rebuilt at the next solve if this changes its levels
    } else if (k == "multigridJacobiWeight") {
      const double w = std::stod(v);
      if (!(w > 0. && w <= 1.5)) fail(PHASE_HIP_EINVAL, "multigridJacobiWeight must be in (0, 1.5]");
      if (w != s->mlJacobiWeight) mlValuesChanged(s);
      s->mlJacobiWeight = w;
    } else if (k == "multigridDensePivoting") {
      const int pv = lv == "always" ? 1 : lv == "auto" ? 0 : -1;
      if (pv < 0) fail(PHASE_HIP_EINVAL, "multigridDensePivoting is AUTO or ALWAYS");
      if (pv != s->mlDensePivot) mlValuesChanged(s);
      s->mlDensePivot = pv;
    } else if (k == "learnedObjective") {
      const int o = lv == "error" ? 0 : lv == "residual" ? 1 : -1;
      if (o < 0) fail(PHASE_HIP_EINVAL, "learnedObjective is ERROR or RESIDUAL");
      s->learnedObjective = o;
    } else if (k == "multigridFinePostSmoothing") {
      s->mlFinePost = (lv == "1" || lv == "true" || lv == "yes") ? 1 : 0;
    } else if (k == "multigridCoarseRows") {
      const int r = std::stoi(v);
      if (r < 2 || r > 64) fail(PHASE_HIP_EINVAL, "multigridCoarseRows must be in [2,64]");
      if (r != s->mlCoarseRows) mlFree(s);
      s->mlCoarseRows = r;
    } else if (k == "multigridDenseRows") {
      const int r = std::stoi(v);
      if (r < 1 || r > 512) fail(PHASE_HIP_EINVAL, "multigridDenseRows must be in [1,512]");
      if (r != s->mlDenseRows) mlFree(s);
      s->mlDenseRows = r;
    } else if (k == "pattern") {
      if (lv == "exact") s->unionPattern = 0;
      else if (lv == "union") s->unionPattern = 1;
      else fail(PHASE_HIP_EINVAL, "unknown pattern policy '" + v + "' (exact or union)");
    }
    // other keys are ignored, as the reference's setup(ptree) reads only the keys it knows
  });
}

int phase_hip_solver_set_rank(phase_hip_solver *s, int rank) {
  return guarded(s, [&] {
    if (rank <= 0) fail(PHASE_HIP_EINVAL, "rank must be positive");
    if (rank != s->rank) {
      s->havePattern = false;
      s->haveRhs = s->haveGuess = s->haveSolution = false;
      s->rank = 0;
      solverFreeBinding(s);
    }
    s->rank = rank;
  });
}

int phase_hip_solver_set(phase_hip_solver *s, int n, const int *rowPtr, const int *colInd,
                         const double *vals) {
  return guarded(s, [&] {
    if (n <= 0 || !rowPtr || !colInd || !vals) fail(PHASE_HIP_EINVAL, "bad arguments to set()");
    if (s->rank && n != s->rank) fail(PHASE_HIP_EINVAL, "set(): size differs from setRank");
    setDevice(s);
    const int nslots = rowPtr[n];
    bool same = s->havePattern && (int)s->patRowPtr.size() == n + 1 &&
                std::memcmp(s->patRowPtr.data(), rowPtr, (n + 1) * sizeof(int)) == 0 &&
                (int)s->patColInd.size() == nslots &&
                std::memcmp(s->patColInd.data(), colInd, nslots * sizeof(int)) == 0;
    if (!same) {
      solverFreeBinding(s);
      s->rank = 0;
      solverSetPattern(s, n, rowPtr, colInd);
    }
    PH_CHECK(hipMemcpyAsync(s->dVals.p, vals, nslots * sizeof(double), hipMemcpyHostToDevice, s->stream));
    solverValuesChanged(s);
  });
}

int phase_hip_solver_set_rhs(phase_hip_solver *s, const double *b) {
  return guarded(s, [&] {
    if (!s->havePattern) fail(PHASE_HIP_ESTATE, "setRhs before set()");
    setDevice(s);
    const int n = s->rank;
    s->dR64.alloc(n);
    PH_CHECK(hipMemcpyAsync(s->dR64.p, b, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL((kPermuteIn<double, double>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                       s->dPerm.p, s->dR64.p, s->dB.p, 1.0);
    PH_CHECK(hipGetLastError());
    s->haveRhs = true;
  });
}

int phase_hip_solver_set_guess(phase_hip_solver *s, const double *x0) {
  return guarded(s, [&] {
    if (!s->havePattern) fail(PHASE_HIP_ESTATE, "setGuess before set()");
    setDevice(s);
    const int n = s->rank;
    s->dR64.alloc(n);
    PH_CHECK(hipMemcpyAsync(s->dR64.p, x0, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL((kPermuteIn<double, double>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                       s->dPerm.p, s->dR64.p, s->dX.p, 1.0);
    PH_CHECK(hipGetLastError());
    s->haveGuess = true;
  });
}

int phase_hip_solver_zero_guess(phase_hip_solver *s) {
  return guarded(s, [&] {
    if (!s->havePattern) fail(PHASE_HIP_ESTATE, "setGuess before set()");
    setDevice(s);
    PH_CHECK(hipMemsetAsync(s->dX.p, 0, s->rank * sizeof(double), s->stream));
    s->haveGuess = true;
  });
}

int phase_hip_malloc(int device, size_t bytes, void **ptr) {
  if (!ptr) return PHASE_HIP_EINVAL;
  *ptr = nullptr;
  if (hipSetDevice(device) != hipSuccess) return PHASE_HIP_EDEVICE;
  return hipMalloc(ptr, bytes ? bytes : 1) == hipSuccess ? PHASE_HIP_OK : PHASE_HIP_EDEVICE;
}

int phase_hip_free(void *ptr) { return hipFree(ptr) == hipSuccess ? PHASE_HIP_OK : PHASE_HIP_EDEVICE; }

int phase_hip_memcpy(void *dst, const void *src, size_t bytes, int toDevice) {
  hipError_t e = hipMemcpy(dst, src, bytes, toDevice ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost);
  return e == hipSuccess ? PHASE_HIP_OK : PHASE_HIP_EDEVICE;
}

int phase_hip_copy_device(void *dst, const void *src, size_t bytes) {
  if ((!dst || !src) && bytes) return PHASE_HIP_EINVAL;
  if (((uintptr_t)dst | (uintptr_t)src | bytes) & 15) return PHASE_HIP_EINVAL;
  if (!bytes) return PHASE_HIP_OK;
  const size_t n16 = bytes / 16;
  const unsigned grid = (unsigned)std::min<size_t>((n16 + 255) / 256, 8192);
  hipLaunchKernelGGL(kCopy16, dim3(grid), dim3(256), 0, 0, n16, (const uint4 *)src, (uint4 *)dst);
  if (hipGetLastError() != hipSuccess) return PHASE_HIP_EDEVICE;
  return hipDeviceSynchronize() == hipSuccess ? PHASE_HIP_OK : PHASE_HIP_EDEVICE;
}

int phase_hip_solver_solve(phase_hip_solver *s, double *relres) {
  return guarded(s, [&] {
    if (!s->havePattern || !s->haveValues) fail(PHASE_HIP_ESTATE, "solve before set()");
    if (!s->haveRhs) fail(PHASE_HIP_ESTATE, "solve before setRhs()");
    setDevice(s);
    checkAssembly(s);
    // Warm start: like the reference's Tpetra x_ (TrilinosSparseMatrixSolver.cpp:31-38),
    // the previous solution is the initial guess unless setGuess() gave one.
    if (!s->haveGuess && !s->haveSolution) PH_CHECK(hipMemsetAsync(s->dX.p, 0, s->rank * sizeof(double), s->stream));
    PH_CHECK(hipEventRecord(s->ev0, s->stream));
    if (s->precond == 1) {
      // hierarchy once per pattern, coarse matrices and factors once per values
      mlEnsureHierarchy(s);
      mlEnsureValues(s, s->mode != PHASE_HIP_MODE_FP64);
      s->mlKCycle = s->mlKCycleSetting >= 0 ? s->mlKCycleSetting : (s->cg || s->gmres ? 1 : 0);
    }
    s->fallbacks = 0;
    s->learnedSteps = 0;
    if (s->gmres) {
      const double b2 = normB2(s);
      s->innerIters = s->nIters = s->outerPasses = 0;
      if (b2 == 0.) {
        PH_CHECK(hipMemsetAsync(s->dX.p, 0, s->rank * sizeof(double), s->stream));
        s->error = 0.;
      } else {
        solveGmres(s, b2);
      }
    } else if (s->mode == PHASE_HIP_MODE_FP64) {
      solveFp64(s);
    } else {
      solveMixed(s);
    }
    PH_CHECK(hipEventRecord(s->ev1, s->stream));
    PH_CHECK(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    PH_CHECK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->solveMs = ms;
    s->haveGuess = false;
    s->haveSolution = true;
    if (relres) *relres = s->error;
    if (!(s->error <= s->tol))
      fail(PHASE_HIP_ENOCONV, "Krylov solve stopped at relative residual " + std::to_string(s->error) +
                                  " after " + std::to_string(s->nIters) + " iterations");
  });
}

int phase_hip_solver_get_x(phase_hip_solver *s, double *x) {
  return guarded(s, [&] {
    if (!s->havePattern) fail(PHASE_HIP_ESTATE, "no system");
    setDevice(s);
    const int n = s->rank;
    s->dR64.alloc(n);
    hipLaunchKernelGGL((kPermuteOut<double>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                       s->dPerm.p, s->dX.p, s->dR64.p);
    PH_CHECK(hipMemcpyAsync(x, s->dR64.p, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    PH_CHECK(hipStreamSynchronize(s->stream));
  });
}

int phase_hip_solver_get_x_device(phase_hip_solver *s, double *x) {
  return guarded(s, [&] {
    if (!s->havePattern) fail(PHASE_HIP_ESTATE, "no system");
    if (!x) fail(PHASE_HIP_EINVAL, "null output");
    setDevice(s);
    const int n = s->rank;
    hipLaunchKernelGGL((kPermuteOut<double>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                       s->dPerm.p, s->dX.p, x);
    PH_CHECK(hipGetLastError());
    PH_CHECK(hipStreamSynchronize(s->stream));
  });
}

int phase_hip_solver_set_guess_device(phase_hip_solver *s, const double *x0) {
  return guarded(s, [&] {
    if (!s->havePattern) fail(PHASE_HIP_ESTATE, "setGuess before set()");
    if (!x0) fail(PHASE_HIP_EINVAL, "null guess");
    setDevice(s);
    const int n = s->rank;
    hipLaunchKernelGGL((kPermuteIn<double, double>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                       s->dPerm.p, x0, s->dX.p, 1.0);
    PH_CHECK(hipGetLastError());
    s->haveGuess = true;
  });
}

int phase_hip_solver_n_iters(const phase_hip_solver *s) { return s ? s->nIters : 0; }
double phase_hip_solver_error(const phase_hip_solver *s) { return s ? s->error : 0.0; }

int phase_hip_solver_stats(const phase_hip_solver *s, int *innerIters, int *outerPasses, int *fallbacks,
                           double *solveMs) {
  if (!s) return PHASE_HIP_EINVAL;
  if (innerIters) *innerIters = s->innerIters;
  if (outerPasses) *outerPasses = s->outerPasses;
  if (fallbacks) *fallbacks = s->fallbacks;
  if (solveMs) *solveMs = s->solveMs;
  return PHASE_HIP_OK;
}

int phase_hip_solver_ordering_info(const phase_hip_solver *s, int *nBlocks, int *nColours) {
  if (!s || !s->havePattern) return PHASE_HIP_ESTATE;
  if (nBlocks) *nBlocks = s->ord.nBlocks;
  if (nColours) *nColours = s->ord.nColours;
  return PHASE_HIP_OK;
}

int phase_hip_solver_multigrid_info(const phase_hip_solver *s, int *nLevels, int *sizes, int cap) {
  if (!s) return PHASE_HIP_EINVAL;
  const std::vector<int> v = mlLevelSizes(s);
  if (nLevels) *nLevels = (int)v.size();
  if (sizes)
    for (int i = 0; i < std::min(cap, (int)v.size()); ++i) sizes[i] = v[i];
  return PHASE_HIP_OK;
}

int phase_hip_solver_multigrid_coarsest(phase_hip_solver *s, int n, double *matrix, double *inverse) {
  return guarded(s, [&] {
    setDevice(s);
    mlCoarsest(s, n, matrix, inverse);
  });
}

int phase_hip_solver_export(phase_hip_solver *s, int *nnz, int *rowPtr, int *colInd, double *vals,
                            double *rhs) {
  return guarded(s, [&] {
    if (!s->havePattern || !s->haveValues) fail(PHASE_HIP_ESTATE, "no system");
    const Ordering &o = s->ord;
    if (nnz) *nnz = o.nnz;
    if (!rowPtr) return;
    setDevice(s);
    std::vector<double> hv(s->dVals.n);
    PH_CHECK(hipMemcpyAsync(hv.data(), s->dVals.p, hv.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    PH_CHECK(hipStreamSynchronize(s->stream));
    std::copy(o.cRowPtr.begin(), o.cRowPtr.end(), rowPtr);
    std::copy(o.cCol.begin(), o.cCol.end(), colInd);
    for (int j = 0; j < o.nnz; ++j) vals[j] = hv[o.compactSrc[j]];
    if (rhs) {
      const int n = s->rank;
      std::vector<double> hb(n, 0.0);
      if (s->haveRhs) {
        s->dR64.alloc(n);
        hipLaunchKernelGGL((kPermuteOut<double>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                           s->dPerm.p, s->dB.p, s->dR64.p);
        PH_CHECK(hipMemcpyAsync(hb.data(), s->dR64.p, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        PH_CHECK(hipStreamSynchronize(s->stream));
      }
      for (int i = 0; i < n; ++i) rhs[i] = -hb[i];
    }
  });
}

int phase_hip_solver_spmv(phase_hip_solver *s, int single, const double *x, double *y) {
  return guarded(s, [&] {
    if (!s->havePattern || !s->haveValues) fail(PHASE_HIP_ESTATE, "no system");
    setDevice(s);
    const int n = s->rank;
    s->dR64.alloc(n);
    PH_CHECK(hipMemcpyAsync(s->dR64.p, x, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
    if (!single) {
      s->k64.alloc(n);
      hipLaunchKernelGGL((kPermuteIn<double, double>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                         s->dPerm.p, s->dR64.p, s->k64.y.p, 1.0);
      hipLaunchKernelGGL((kSpmv<double>), dim3(spmvGrid(s)), dim3(kRedBlock), 0, s->stream, n, s->ord.nSlices,
                         s->dSliceOff.p, s->dSellCol.p, s->dSell64.p, s->k64.y.p, s->k64.t.p);
      hipLaunchKernelGGL((kPermuteOut<double>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                         s->dPerm.p, s->k64.t.p, s->dR64.p);
    } else {
      ensureSingle(s);
      s->k32.alloc(n);
      hipLaunchKernelGGL((kPermuteIn<double, float>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                         s->dPerm.p, s->dR64.p, s->k32.y.p, 1.0);
      hipLaunchKernelGGL((kSpmv<float>), dim3(spmvGrid(s)), dim3(kRedBlock), 0, s->stream, n, s->ord.nSlices,
                         s->dSliceOff.p, s->dSellCol.p, s->dSell32.p, s->k32.y.p, s->k32.t.p);
      hipLaunchKernelGGL((kPermuteOut<float>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                         s->dPerm.p, s->k32.t.p, s->dR64.p);
    }
    PH_CHECK(hipGetLastError());
    PH_CHECK(hipMemcpyAsync(y, s->dR64.p, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    PH_CHECK(hipStreamSynchronize(s->stream));
  });
}

int phase_hip_solver_spmv_noalloc(phase_hip_solver *s, int single, const double *x, double *y) {
  return guarded(s, [&] {
    if (s->rank > kNoAllocMaxRank)
      fail(PHASE_HIP_EINVAL, "rank " + std::to_string(s->rank) + " exceeds the allocation-free scratch (" +
                                 std::to_string(kNoAllocMaxRank) + " rows)");
    if (!s->havePattern || !s->haveValues) fail(PHASE_HIP_ESTATE, "no system");
    if (single && !s->sell32Valid)
      fail(PHASE_HIP_ESTATE, "no FP32 copy of the matrix yet (a mixed solve or an FP32 spmv builds it)");
    if (!x || !y) fail(PHASE_HIP_EINVAL, "null vector");
    setDevice(s);
    double *sp = gScratchPtr[s->device];
    if (!sp) fail(PHASE_HIP_ESTATE, "scratch not set up on this device");
    const int n = s->rank;
    // xs: operand in solver order; ys: first the operand in the caller's order,
    // then the product; the product in the caller's order goes back to xs
    double *xs = sp, *ys = sp + kNoAllocMaxRank;
    PH_CHECK(hipMemcpyAsync(ys, x, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
    if (!single) {
      hipLaunchKernelGGL((kPermuteIn<double, double>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                         s->dPerm.p, ys, xs, 1.0);
      hipLaunchKernelGGL((kSpmv<double>), dim3(spmvGrid(s)), dim3(kRedBlock), 0, s->stream, n, s->ord.nSlices,
                         s->dSliceOff.p, s->dSellCol.p, s->dSell64.p, xs, ys);
      hipLaunchKernelGGL((kPermuteOut<double>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                         s->dPerm.p, ys, xs);
    } else {
      float *xf = reinterpret_cast<float *>(xs), *yf = reinterpret_cast<float *>(ys);
      hipLaunchKernelGGL((kPermuteIn<double, float>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                         s->dPerm.p, ys, xf, 1.0);
      hipLaunchKernelGGL((kSpmv<float>), dim3(spmvGrid(s)), dim3(kRedBlock), 0, s->stream, n, s->ord.nSlices,
                         s->dSliceOff.p, s->dSellCol.p, s->dSell32.p, xf, yf);
      hipLaunchKernelGGL((kPermuteOut<float>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                         s->dPerm.p, yf, xs);
    }
    PH_CHECK(hipGetLastError());
    PH_CHECK(hipMemcpyAsync(y, xs, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    PH_CHECK(hipStreamSynchronize(s->stream));
  });
}

int phase_hip_solver_ilu_export(phase_hip_solver *s, int *nnz, int *perm, int *rowPtr, int *colInd,
                                double *vals) {
  return guarded(s, [&] {
    if (!s->havePattern || !s->haveValues) fail(PHASE_HIP_ESTATE, "no system");
    setDevice(s);
    const Ordering &o = s->ord;
    const int n = s->rank;
    long total = 0;
    for (int k = 0; k < n; ++k) total += o.rows[k].nL + o.rows[k].nU + 1;
    if (nnz) *nnz = (int)total;
    if (!rowPtr) return;
    ensureFactor64(s);
    std::vector<double> f(s->dFac64.n);
    PH_CHECK(hipMemcpyAsync(f.data(), s->dFac64.p, f.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    PH_CHECK(hipStreamSynchronize(s->stream));
    std::copy(o.perm.begin(), o.perm.end(), perm);
    int w = 0;
    rowPtr[0] = 0;
    for (int b = 0; b < o.nBlocks; ++b) {
      const BlockInfo &B = o.blocks[b];
      for (int lane = 0; lane < B.nRows; ++lane) {
        const int k = B.rowStart + lane;
        const RowInfo &ri = o.rows[k];
        for (int e = 0; e < ri.nL; ++e) {
          int sl = B.lOff + e * kLanes + lane;
          colInd[w] = o.lCol[sl];
          vals[w++] = f[sl];
        }
        colInd[w] = k;
        vals[w++] = f[(size_t)o.nLSlots + o.nUSlots + k];
        for (int e = 0; e < ri.nU; ++e) {
          int sl = B.uOff + e * kLanes + lane;
          colInd[w] = o.uCol[sl];
          vals[w++] = f[(size_t)o.nLSlots + sl];
        }
        rowPtr[k + 1] = w;
      }
    }
  });
}

}  // extern "C"

namespace {

// Algorithmic bytes of the timed launches (entries as phase_hip_solver_time_kernels):
// every matrix or factor entry touched (value + column index), every vector
// element read or written once, one gathered value per off-block entry, and
// the block / lane records the sweeps read.
void kernelBytes(const Ordering &o, double *out) {
  const double n = o.n;
  const int C = (int)o.colourPtr.size() - 1;
  double nnzL = 0., nnzU = 0.;
  for (const RowInfo &r : o.rows) {
    nnzL += r.nL;
    nnzU += r.nU;
  }
  // forward / backward sweep over blocks [b0, b1) with value size sz
  auto sweep = [&](int b0, int b1, double sz, bool upper) {
    double rows = 0., ent = 0., off = 0.;
    for (int b = b0; b < b1; ++b) {
      const BlockInfo &B = o.blocks[b];
      rows += B.nRows;
      for (int k = B.rowStart; k < B.rowStart + B.nRows; ++k) {
        const RowInfo &r = o.rows[k];
        ent += upper ? r.nU : r.nL;
        off += upper ? r.nU - r.nUin : r.nLoff;
      }
    }
    const double recs = (b1 - b0) * (sizeof(BlockInfo) + 64.0 * sizeof(LaneRec));
    return ent * (sz + 4) + off * sz + rows * sz * (upper ? 3 : 2) + recs;
  };
  auto spmv = [&](double sz) { return o.nnz * (sz + 4) + 2 * n * sz + (o.nSlices + 1) * 4.0; };
  auto factor = [&](double sz) {
    return (nnzL + nnzU + n) * sz * 2 + nnzL * 4 + (o.updSrc.size() + o.updDst.size()) * 4.0 +
           (o.updPtr.size()) * 4.0 + o.facSrc.size() * 4.0;
  };
  out[0] = spmv(8);
  out[1] = spmv(4);
  out[2] = sweep(0, o.nBlocks, 8, false);
  out[3] = sweep(0, o.nBlocks, 8, true);
  out[4] = sweep(0, o.nBlocks, 4, false);
  out[5] = sweep(0, o.nBlocks, 4, true);
  out[6] = factor(8);
  out[7] = factor(4);
  out[11] = out[0];
  out[12] = out[1];
  // per-launch sweeps, FP64 in 8..10 and FP32 in 13..15
  for (const int sz : {8, 4}) {
    double *o3 = sz == 8 ? out + 8 : out + 13;  // [U launch, fused LU launch, L launch]
    o3[0] = o3[1] = o3[2] = 0.;
    if (C >= 2) {
      o3[0] = sweep(0, o.colourPtr[C - 1], sz, true) / (C - 1);
      o3[2] = sweep(0, o.colourPtr[C - 1], sz, false) / (C - 1);
    }
    if (C >= 1) {
      // fused: records read once, y kept in registers (neither stored nor reloaded)
      const int b0 = o.colourPtr[C - 1];
      double rows = 0.;
      for (int b = b0; b < o.nBlocks; ++b) rows += o.blocks[b].nRows;
      o3[1] = sweep(b0, o.nBlocks, sz, false) + sweep(b0, o.nBlocks, sz, true) -
              (o.nBlocks - b0) * (sizeof(BlockInfo) + 64.0 * sizeof(LaneRec)) - 2 * rows * sz;
    }
    // the CG forms (prologue kProCgR: r read and written, q read, x read and
    // written in FP64, on top of the plain sweep), FP64 in 16..17, FP32 in 18..19:
    // [L launch, fused LU launch]
    double *oc = sz == 8 ? out + 16 : out + 18;
    oc[0] = oc[1] = 0.;
    if (C >= 2) {
      double rows = 0.;
      for (int b = 0; b < o.colourPtr[C - 1]; ++b) rows += o.blocks[b].nRows;
      oc[0] = o3[2] + rows / (C - 1) * (3 * sz + 16);
    }
    if (C >= 1) {
      double rows = 0.;
      for (int b = o.colourPtr[C - 1]; b < o.nBlocks; ++b) rows += o.blocks[b].nRows;
      oc[1] = o3[1] + rows * (3 * sz + 16);
    }
  }
  // multigrid residual t = b - A x: the FP64 matrix, x gathered, b read, t written
  out[20] = o.nnz * 12.0 + 3 * n * 4 + (o.nSlices + 1) * 4.0;
  out[21] = o.nnz * 12.0 + 3 * n * 8 + (o.nSlices + 1) * 4.0;
}

}  // namespace

extern "C" {

int phase_hip_solver_time_kernels(phase_hip_solver *s, int reps, double *ms, int nms) {
  return guarded(s, [&] {
    if (!s->havePattern || !s->haveValues) fail(PHASE_HIP_ESTATE, "no system");
    if (!ms || nms < 1) fail(PHASE_HIP_EINVAL, "no output");
    setDevice(s);
    const int n = s->rank;
    reps = std::max(1, reps);
    ensureFactor64(s);
    ensureSingle(s);
    s->k64.alloc(n);
    s->k32.alloc(n);
    hipLaunchKernelGGL((kConvert<double, double>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                       s->dB.p, s->k64.ph.p);
    hipLaunchKernelGGL((kConvert<double, float>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                       s->dB.p, s->k32.ph.p);
    auto timeIt = [&](auto &&launch) {
      launch();  // warm
      PH_CHECK(hipEventRecord(s->ev0, s->stream));
      for (int r = 0; r < reps; ++r) launch();
      PH_CHECK(hipEventRecord(s->ev1, s->stream));
      PH_CHECK(hipEventSynchronize(s->ev1));
      float t = 0.f;
      PH_CHECK(hipEventElapsedTime(&t, s->ev0, s->ev1));
      return (double)t / reps;
    };
    const Scalars *noSc = idleScalars(s);
    auto lsweep = [&](auto *fac, auto *in, auto *out) {
      using T = std::remove_pointer_t<decltype(in)>;
      launchL<T, kProNone>(s, iluDev<T>(s, fac), noSc, (const T *)in, out, (T *)nullptr, (const T *)nullptr,
                           (const T *)nullptr);
    };
    auto usweep = [&](auto *fac, auto *in, auto *out) {
      using T = std::remove_pointer_t<decltype(in)>;
      launchU<T>(s, iluDev<T>(s, fac), noSc, (const T *)in, out);
    };
    double out[kTimedKernels];
    out[0] = timeIt([&] {
      hipLaunchKernelGGL((kSpmv<double>), dim3(spmvGrid(s)), dim3(kRedBlock), 0, s->stream, n, s->ord.nSlices,
                         s->dSliceOff.p, s->dSellCol.p, s->dSell64.p, s->k64.ph.p, s->k64.t.p);
    });
    out[1] = timeIt([&] {
      hipLaunchKernelGGL((kSpmv<float>), dim3(spmvGrid(s)), dim3(kRedBlock), 0, s->stream, n, s->ord.nSlices,
                         s->dSliceOff.p, s->dSellCol.p, s->dSell32.p, s->k32.ph.p, s->k32.t.p);
    });
    out[2] = timeIt([&] { lsweep(s->dFac64.p, s->k64.ph.p, s->k64.y.p); });
    out[3] = timeIt([&] { usweep(s->dFac64.p, s->k64.y.p, s->k64.sh.p); });
    out[4] = timeIt([&] { lsweep(s->dFac32.p, s->k32.ph.p, s->k32.y.p); });
    out[5] = timeIt([&] { usweep(s->dFac32.p, s->k32.y.p, s->k32.sh.p); });
    out[6] = timeIt([&] { factorize<double>(s, s->dFac64); });
    out[7] = timeIt([&] { factorize<float>(s, s->dFac32); });
    s->fac32Valid = false;  // solves use the rounded FP64 factor
    // the launches of one preconditioner apply inside a solve: kUSolve and
    // kLSolve over every colour but the last, kLUSolve over the last; FP64
    // in 8..10, FP32 in 13..15
    const std::vector<int> &cp = s->ord.colourPtr;
    const int C = (int)cp.size() - 1;
    s->dR64.alloc(n);
    PH_CHECK(hipMemcpyAsync(s->dR64.p, s->dB.p, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    auto perLaunch = [&](auto *fac, auto &K, double *o3, double *oc) {
      using T = std::remove_pointer_t<decltype(fac)>;
      const IluDev<T> L = iluDev<T>(s, fac);
      const T *in = K.ph.p;
      T *y = K.y.p, *x = K.sh.p;
      o3[0] = o3[1] = o3[2] = 0.;
      oc[0] = oc[1] = 0.;
      // the CG forms read p, r, q and x (alpha = 0 on the idle scalars: nothing moves)
      for (T *v : {K.p.p, K.r.p, K.v.p})
        hipLaunchKernelGGL((kConvert<double, T>), dim3(gridFor(n, 256, kElemCap)), dim3(256), 0, s->stream, n,
                           s->dB.p, v);
      byWidth(s, [&](auto Wc) {
        constexpr int W = decltype(Wc)::value;
        if (C >= 2) {
          o3[0] = timeIt([&] {
                    for (int c = C - 2; c >= 0; --c)
                      hipLaunchKernelGGL((kUSolve<T, W>), dim3(cp[c + 1] - cp[c]), dim3(64), 0, s->stream, L,
                                         cp[c], noSc, (const T *)y, x, (const T *)nullptr, (T *)nullptr);
                  }) / (C - 1);
          o3[2] = timeIt([&] {
                    for (int c = 0; c + 1 < C; ++c)
                      hipLaunchKernelGGL((kLSolve<T, kProNone, W>), dim3(cp[c + 1] - cp[c]), dim3(64), 0,
                                         s->stream, L, cp[c], noSc, in, y, (T *)nullptr, (const T *)nullptr,
                                         (const T *)nullptr, (double *)nullptr);
                  }) / (C - 1);
        }
        if (C >= 1)
          o3[1] = timeIt([&] {
            hipLaunchKernelGGL((kLUSolve<T, kProNone, W>), dim3(cp[C] - cp[C - 1]), dim3(64), 0, s->stream, L,
                               cp[C - 1], noSc, in, (const T *)y, (T *)nullptr, (const T *)nullptr,
                               (const T *)nullptr, x, (double *)nullptr, (const T *)nullptr, (T *)nullptr);
          });
        if (C >= 2)
          oc[0] = timeIt([&] {
                    for (int c = 0; c + 1 < C; ++c)
                      hipLaunchKernelGGL((kLSolve<T, kProCgR, W>), dim3(cp[c + 1] - cp[c]), dim3(64), 0,
                                         s->stream, L, cp[c], noSc, (const T *)K.p.p, y, K.r.p, (const T *)nullptr,
                                         (const T *)K.v.p, s->dR64.p);
                  }) / (C - 1);
        if (C >= 1)
          oc[1] = timeIt([&] {
            hipLaunchKernelGGL((kLUSolve<T, kProCgR, W>), dim3(cp[C] - cp[C - 1]), dim3(64), 0, s->stream, L,
                               cp[C - 1], noSc, (const T *)K.p.p, (const T *)y, K.r.p, (const T *)nullptr,
                               (const T *)K.v.p, x, s->dR64.p, (const T *)nullptr, (T *)nullptr);
          });
      });
    };
    perLaunch(s->dFac64.p, s->k64, out + 8, out + 16);
    ensureSingle(s);  // the FP32 copy of the FP64 factor, as mixed solves use it
    perLaunch(s->dFac32.p, s->k32, out + 13, out + 18);
    // the multigrid cycle's fine-level residual t = b - A x (FP64 values), FP32 in 20, FP64 in 21
    out[20] = timeIt([&] { mlResidLaunch<float>(s, s->k32.ph.p, s->k32.p.p, s->k32.t.p); });
    out[21] = timeIt([&] { mlResidLaunch<double>(s, s->k64.ph.p, s->k64.p.p, s->k64.t.p); });
    // cache-cold SpMV: before each timed launch a 1 GiB copy evicts the matrix
    // and vectors from L2 and from the 256 MB last-level cache, and the events
    // bracket the one launch
    {
      const size_t flushBytes = (size_t)1 << 30;
      void *flush = nullptr;
      PH_CHECK(hipMalloc(&flush, flushBytes));
      const size_t n16 = flushBytes / 32;
      auto cold = [&](auto &&launch) {
        double tot = 0.;
        for (int r = 0; r < reps; ++r) {
          hipLaunchKernelGGL(kCopy16, dim3(8192), dim3(256), 0, s->stream, n16, (const uint4 *)flush,
                             (uint4 *)flush + n16);
          PH_CHECK(hipEventRecord(s->ev0, s->stream));
          launch();
          PH_CHECK(hipEventRecord(s->ev1, s->stream));
          PH_CHECK(hipEventSynchronize(s->ev1));
          float t = 0.f;
          PH_CHECK(hipEventElapsedTime(&t, s->ev0, s->ev1));
          tot += t;
        }
        return tot / reps;
      };
      out[11] = cold([&] {
        hipLaunchKernelGGL((kSpmv<double>), dim3(spmvGrid(s)), dim3(kRedBlock), 0, s->stream, n, s->ord.nSlices,
                         s->dSliceOff.p, s->dSellCol.p, s->dSell64.p, s->k64.ph.p, s->k64.t.p);
      });
      out[12] = cold([&] {
        hipLaunchKernelGGL((kSpmv<float>), dim3(spmvGrid(s)), dim3(kRedBlock), 0, s->stream, n, s->ord.nSlices,
                         s->dSliceOff.p, s->dSellCol.p, s->dSell32.p, s->k32.ph.p, s->k32.t.p);
      });
      PH_CHECK(hipFree(flush));
    }
    checkDeviceErr(s);
    for (int i = 0; i < std::min(nms, kTimedKernels); ++i) ms[i] = out[i];
  });
}

int phase_hip_solver_kernel_bytes(phase_hip_solver *s, double *bytes, int nb) {
  return guarded(s, [&] {
    if (!s->havePattern) fail(PHASE_HIP_ESTATE, "no pattern");
    if (!bytes || nb < 1) fail(PHASE_HIP_EINVAL, "no output");
    double out[kTimedKernels];
    kernelBytes(s->ord, out);
    for (int i = 0; i < std::min(nb, kTimedKernels); ++i) bytes[i] = out[i];
  });
}

}  // extern "C"
