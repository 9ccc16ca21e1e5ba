// Launchers shared by the solver and the multigrid hierarchy (included by
// solver.hip and multilevel.hip only): grid sizes, the ILU(0) sweeps and
// factorisation of one ordering, and the lazily built FP64 / FP32 factors.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "solver.h"
#include "solver_kernels.h"

namespace phase {

inline int gridFor(long n, int block, int cap) {
  long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return (int)std::min<long>(g, cap);
}

constexpr int kElemCap = 4096;
constexpr int kRedCap = 1024;


// entries of phase_hip_solver_time_kernels / phase_hip_solver_kernel_bytes
constexpr int kTimedKernels = 22;

template <class T>
IluDev<T> iluDev(phase_hip_solver *s, T *fac) {
  IluDev<T> L;
  L.blocks = s->dBlocks.p;
  L.rows = s->dRows.p;
  L.lanes = s->dLanes.p;
  L.lCol = s->dLCol.p;
  L.uCol = s->dUCol.p;
  L.fac = fac;
  L.nb = s->ord.nBlocks;
  L.nLSlots = s->ord.nLSlots;
  L.nUSlots = s->ord.nUSlots;
  L.n = s->rank;
  return L;
}

// Scalars slot that is never set done: sweeps run outside a Krylov iteration
inline const Scalars *idleScalars(phase_hip_solver *s) { return s->dSc.p + 3; }

// Calls f with the ordering's sweep register width as a compile-time constant.
template <class F>
void byWidth(const phase_hip_solver *s, F &&f) {
  if (s->ord.regWidth == 4) f(std::integral_constant<int, 4>());
  else f(std::integral_constant<int, 8>());
}

// Sweep launchers: one launch per block colour (forward colours ascending,
// backward descending).
template <class T, int PRO>
void launchL(phase_hip_solver *s, const IluDev<T> &L, const Scalars *sc, const T *in, T *y, T *pv, const T *r,
             const T *v, double *xs = nullptr) {
  byWidth(s, [&](auto Wc) {
    constexpr int W = decltype(Wc)::value;
    const std::vector<int> &cp = s->ord.colourPtr;
    for (int c = 0; c + 1 < (int)cp.size(); ++c)
      if (cp[c + 1] > cp[c])
        hipLaunchKernelGGL((kLSolve<T, PRO, W>), dim3(cp[c + 1] - cp[c]), dim3(64), 0, s->stream, L, cp[c], sc, in,
                           y, pv, r, v, xs);
  });
}

template <class T>
void launchU(phase_hip_solver *s, const IluDev<T> &L, const Scalars *sc, const T *y, T *x) {
  byWidth(s, [&](auto Wc) {
    constexpr int W = decltype(Wc)::value;
    const std::vector<int> &cp = s->ord.colourPtr;
    for (int c = (int)cp.size() - 2; c >= 0; --c)
      if (cp[c + 1] > cp[c])
        hipLaunchKernelGGL((kUSolve<T, W>), dim3(cp[c + 1] - cp[c]), dim3(64), 0, s->stream, L, cp[c], sc, y, x,
                           (const T *)nullptr, (T *)nullptr);
  });
}

// out = U^{-1} L^{-1} (right-hand side formed by PRO); the last forward colour
// and the first backward colour run fused in one launch.  With addTo set,
// addTo = addFrom + out as well (row by row, in the launch that finishes it).
template <class T, int PRO>
void launchLU(phase_hip_solver *s, const IluDev<T> &L, const Scalars *sc, const T *in, T *y, T *pv, const T *r,
              const T *v, T *out, double *xs = nullptr, const T *addFrom = nullptr, T *addTo = nullptr) {
  const std::vector<int> &cp = s->ord.colourPtr;
  const int C = (int)cp.size() - 1;
  if (C < 1) return;
  byWidth(s, [&](auto Wc) {
    constexpr int W = decltype(Wc)::value;
    for (int c = 0; c + 1 < C; ++c)
      if (cp[c + 1] > cp[c])
        hipLaunchKernelGGL((kLSolve<T, PRO, W>), dim3(cp[c + 1] - cp[c]), dim3(64), 0, s->stream, L, cp[c], sc, in,
                           y, pv, r, v, xs);
    if (cp[C] > cp[C - 1])
      hipLaunchKernelGGL((kLUSolve<T, PRO, W>), dim3(cp[C] - cp[C - 1]), dim3(64), 0, s->stream, L, cp[C - 1], sc,
                         in, (const T *)y, pv, r, v, out, xs, addFrom, addTo);
    for (int c = C - 2; c >= 0; --c)
      if (cp[c + 1] > cp[c])
        hipLaunchKernelGGL((kUSolve<T, W>), dim3(cp[c + 1] - cp[c]), dim3(64), 0, s->stream, L, cp[c], sc,
                           (const T *)y, out, addFrom, addTo);
  });
}

template <class T>
void launchFactor(phase_hip_solver *s, const IluDev<T> &L) {
  const std::vector<int> &cp = s->ord.colourPtr;
  for (int c = 0; c + 1 < (int)cp.size(); ++c)
    if (cp[c + 1] > cp[c])
      hipLaunchKernelGGL((kIluFactor<T>), dim3(cp[c + 1] - cp[c]), dim3(64), 0, s->stream, L, s->dUpdPtr.p,
                         s->dUpdSrc.p, s->dUpdDst.p, s->dErr.p, cp[c]);
}

inline int spmvGrid(phase_hip_solver *s) { return std::max(1, (s->ord.nSlices + 3) / 4); }


// The three scalar records (Krylov run, outer refinement, misc) to the host.
inline void pullScalars(phase_hip_solver *s) {
  PH_CHECK(hipMemcpyAsync(s->hSc, s->dSc.p, 3 * sizeof(Scalars), hipMemcpyDeviceToHost, s->stream));
  PH_CHECK(hipStreamSynchronize(s->stream));
}

inline void checkDeviceErr(phase_hip_solver *s) {
  int e = 0;
  PH_CHECK(hipMemcpyAsync(&e, s->dErr.p, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  PH_CHECK(hipStreamSynchronize(s->stream));
  if (e & 2) {
    PH_CHECK(hipMemsetAsync(s->dErr.p, 0, sizeof(int), s->stream));
    fail(PHASE_HIP_EBREAKDOWN, "zero pivot in ILU(0) factorisation");
  }
}

// fac = [L slots | U slots | diagonal (n) | inverse diagonal (n)]
inline size_t facSize(const phase_hip_solver *s) { return (size_t)s->ord.nLSlots + s->ord.nUSlots + 2 * (size_t)s->rank; }

template <class T>
void factorize(phase_hip_solver *s, DBuf<T> &fac) {
  const int m = (int)(fac.n - s->rank);  // gathered entries: L, U, diagonal
  hipLaunchKernelGGL((kGather<T>), dim3(gridFor(m, 256, kElemCap)), dim3(256), 0, s->stream, m,
                     s->dFacSrc.p, s->dVals.p, fac.p);
  launchFactor<T>(s, iluDev<T>(s, fac.p));
  hipLaunchKernelGGL((kInvDiag<T>), dim3(gridFor(s->rank, 256, kElemCap)), dim3(256), 0, s->stream, s->rank, fac.p,
                     s->ord.nLSlots + s->ord.nUSlots);
  PH_CHECK(hipGetLastError());
}

inline void ensureFactor64(phase_hip_solver *s) {
  if (s->fac64Valid) return;
  s->dFac64.alloc(facSize(s));
  factorize<double>(s, s->dFac64);
  checkDeviceErr(s);
  s->fac64Valid = true;
}

inline void ensureSingle(phase_hip_solver *s) {
  if (!s->sell32Valid) {
    s->dSell32.alloc(s->dSell64.n);
    const int m = (int)s->dSell64.n;
    hipLaunchKernelGGL((kConvert<double, float>), dim3(gridFor(m, 256, kElemCap)), dim3(256), 0, s->stream,
                       m, s->dSell64.p, s->dSell32.p);
    s->sell32Valid = true;
  }
  if (!s->fac32Valid) {
    // factor in FP64, store and apply in FP32: rounding the finished factor
    // keeps the FP64 pivots, where an FP32 factorisation of an ill-conditioned
    // matrix loses them to cancellation
    ensureFactor64(s);
    const int m = (int)s->dFac64.n;
    s->dFac32.alloc(m);
    hipLaunchKernelGGL((kConvert<double, float>), dim3(gridFor(m, 256, kElemCap)), dim3(256), 0, s->stream, m,
                       s->dFac64.p, s->dFac32.p);
    PH_CHECK(hipGetLastError());
    s->fac32Valid = true;
  }
}


}  // namespace phase
