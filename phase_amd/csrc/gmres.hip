// Flexible restarted GMRES (setup "solver" = "GMRES"): the Krylov method the
// reference's MueLu backend asks Belos for by default
// (src/Math/TrilinosMueluSparseMatrixSolver.cpp:47), for systems that are not
// symmetric (the ghost-cell pressure rows of an immersed boundary).  The
// preconditioner (ILU(0), or the multigrid cycle, whose K cycles make it vary
// from one application to the next) is applied on the right and each
// preconditioned vector is kept, as in Saad's FGMRES (SIAM J. Sci. Comput.
// 14, 1993):
//
//   r = b - A x, beta = ||r||, v_0 = r / beta, g = beta e_1
//   for j = 0 .. m-1:
//     z_j = M^-1 v_j;  w = A z_j
//     h_ij = (w, v_i), w -= sum h_ij v_i   (classical Gram-Schmidt, twice)
//     h_j+1,j = ||w||, v_j+1 = w / h_j+1,j
//     Givens rotations of column j; |g_j+1| is the residual norm
//   y = H^-1 g (upper triangular), x += sum y_j z_j; restart from x
//
// In mode fp64 everything is FP64.  In mode mixed the basis V and the
// preconditioned vectors Z are FP32, and so are the preconditioner (the FP32
// copy of the ILU(0) factor, the multigrid cycle in FP32) and the products'
// vectors (FP64 matrix values, rows accumulated in FP64); the dot products,
// the Hessenberg matrix, the rotations and the solution stay FP64.  Every cycle
// starts from the FP64 true residual, so restarts are the refinement points:
// a mixed cycle stops once its estimate has dropped by kGmMixedDrop (or meets
// the tolerance).  When kGmMixedStall cycles in a row either make no new best
// true residual or leave it well above their own estimate without halving it
// (the FP32 basis no longer carries the correction), or a cycle breaks down
// at its first step, the solve continues in FP64 from the best iterate.
// The Hessenberg column, the rotations and the residual estimate live on the
// device, so an Arnoldi step needs no host round trip; the host checks the
// step count and the convergence flag after every chunk of steps, and every
// launch after convergence returns at once (except the SpMV).  A cycle that
// ends on the estimate is followed by the true residual, which decides.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "multilevel.h"
#include "solver.h"
#include "solver_kernels.h"
#include "solver_launch.h"

namespace phase {

namespace {

constexpr int kGmGroup = 8;
constexpr double kGmMixedDrop = 1e-5;  // a mixed cycle's reduction of the residual estimate
constexpr int kGmMixedStall = 3;

// partials of (w, V_j) for the kGmGroup vectors j = 8 blockIdx.y + q < k
template <class T>
__global__ __launch_bounds__(kRedBlock) void kGmDots(int n, int k, const T *__restrict__ V,
                                                     const T *__restrict__ w, double *partials,
                                                     const Scalars *sc) {
  if (sc->done) return;
  const int g0 = blockIdx.y * kGmGroup;
  double d[kGmGroup];
#pragma unroll
  for (int q = 0; q < kGmGroup; ++q) d[q] = 0.;
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock) {
    const double wi = (double)w[i];
#pragma unroll
    for (int q = 0; q < kGmGroup; ++q)
      if (g0 + q < k) d[q] = fma((double)V[(size_t)(g0 + q) * n + i], wi, d[q]);
  }
  blockPartials<kGmGroup>(d, partials + (size_t)g0 * gridDim.x, blockIdx.x);
}

// h_j = sum of the G partials of vector j (one wavefront per j, fixed order);
// the Hessenberg column takes h (first pass) or adds it (second pass)
__global__ __launch_bounds__(kFinBlock) void kGmDotsFinish(int k, int G, const double *partials, double *hcur,
                                                           double *Hcol, int accumulate, const Scalars *sc) {
  if (sc->done) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j = wave; j < k; j += kFinBlock / 64) {
    double acc = 0.;
    for (int b = lane; b < G; b += 64) acc += partials[(size_t)j * G + b];
    acc = waveSum(acc);
    if (lane == 0) {
      hcur[j] = acc;
      Hcol[j] = accumulate ? Hcol[j] + acc : acc;
    }
  }
}

// w -= sum_{j<k} h_j V_j (in FP64, stored in T); with NORM the partials of
// (w, w) of the stored w
template <class T, bool NORM>
__global__ __launch_bounds__(kRedBlock) void kGmAxpy(int n, int k, const T *__restrict__ V,
                                                     const double *__restrict__ hcur, T *__restrict__ w,
                                                     double *partials, const Scalars *sc) {
  if (sc->done) return;
  __shared__ double h[kGmMaxRestart + 1];
  for (int j = threadIdx.x; j < k; j += kRedBlock) h[j] = hcur[j];
  __syncthreads();
  double d[1] = {0.};
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock) {
    double v = (double)w[i];
    for (int j = 0; j < k; ++j) v = fma(-h[j], (double)V[(size_t)j * n + i], v);
    const T vt = (T)v;
    w[i] = vt;
    if (NORM) d[0] = fma((double)vt, (double)vt, d[0]);
  }
  if (NORM) blockPartials<1>(d, partials, blockIdx.x);
}

// h_j+1,j = ||w||; the earlier rotations and a new one applied to column j;
// g updated; the step count, the residual estimate and the flag published
__global__ __launch_bounds__(kFinBlock) void kGmGivens(int j, int ld, const double *partials, int G, double *H,
                                                       double *cs, double *sn, double *g, double *vscale,
                                                       Scalars *sc) {
  if (sc->done) return;
  double d[1];
  finishPartials<1, kFinBlock>(partials, G, d);
  if (threadIdx.x != 0) return;
  double *h = H + (size_t)j * ld;
  const double hn = sqrt(d[0]);
  for (int i = 0; i < j; ++i) {
    const double a = h[i], b = h[i + 1];
    h[i] = cs[i] * a + sn[i] * b;
    h[i + 1] = -sn[i] * a + cs[i] * b;
  }
  const double a = h[j], b = hn;
  const double r = hypot(a, b);
  const double c = r > 0. ? a / r : 1., s = r > 0. ? b / r : 0.;
  cs[j] = c;
  sn[j] = s;
  h[j] = r;
  h[j + 1] = 0.;
  const double gj = g[j];
  g[j] = c * gj;
  g[j + 1] = -s * gj;
  *vscale = hn > 0. ? 1. / hn : 0.;
  sc->iters = j + 1;
  sc->rr = g[j + 1] * g[j + 1];
  if (!isfinite(r) || !isfinite(g[j + 1])) sc->done = kDoneBreakdown;
  else if (sc->rr <= sc->tol2 || hn == 0.) sc->done = kDoneConverged;
}

// v *= *scale
template <class T>
__global__ void kGmScale(int n, T *__restrict__ v, const double *scale, const Scalars *sc) {
  if (sc->done) return;
  const double f = *scale;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = (T)((double)v[i] * f);
}

// v_0 = r * *scale (the cycle's first basis vector from the FP64 residual)
template <class T>
__global__ void kGmFirst(int n, const double *__restrict__ r, const double *scale, T *__restrict__ v) {
  const double f = *scale;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = (T)(r[i] * f);
}

__global__ void kGmSetTol(double tol2, Scalars *sc) {
  if (threadIdx.x == 0) sc->tol2 = tol2;
}

// start of a cycle from the true residual in sc->rr (v_0 holds r)
__global__ void kGmStart(int m, double *g, double *vscale, Scalars *sc) {
  if (threadIdx.x != 0) return;
  const double beta = sqrt(sc->rr);
  for (int i = 0; i <= m; ++i) g[i] = 0.;
  g[0] = beta;
  *vscale = beta > 0. ? 1. / beta : 0.;
  sc->iters = 0;
  sc->done = !isfinite(sc->rr) ? kDoneBreakdown : sc->rr <= sc->tol2 ? kDoneConverged : kDoneRunning;
}

// y = H(0:k, 0:k)^-1 g(0:k), back substitution (runs whatever sc->done says)
__global__ void kGmSolveY(int k, int ld, const double *H, const double *g, double *y) {
  if (threadIdx.x != 0) return;
  for (int i = k - 1; i >= 0; --i) {
    double t = g[i];
    for (int j = i + 1; j < k; ++j) t -= H[(size_t)j * ld + i] * y[j];
    const double hii = H[(size_t)i * ld + i];
    y[i] = hii != 0. ? t / hii : 0.;
  }
}

// x += sum_{j<k} y_j Z_j (x in FP64)
template <class T>
__global__ __launch_bounds__(kRedBlock) void kGmUpdateX(int n, int k, const T *__restrict__ Z,
                                                        const double *__restrict__ y, double *__restrict__ x) {
  __shared__ double sy[kGmMaxRestart];
  for (int j = threadIdx.x; j < k; j += kRedBlock) sy[j] = y[j];
  __syncthreads();
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock) {
    double v = x[i];
    for (int j = 0; j < k; ++j) v = fma(sy[j], (double)Z[(size_t)j * n + i], v);
    x[i] = v;
  }
}

// The Krylov vectors of precision T
template <class T>
DBuf<T> &gmBasis(phase_hip_solver *s);
template <>
DBuf<double> &gmBasis<double>(phase_hip_solver *s) {
  return s->gmV;
}
template <>
DBuf<float> &gmBasis<float>(phase_hip_solver *s) {
  return s->gmV32;
}
template <class T>
DBuf<T> &gmPrec(phase_hip_solver *s);
template <>
DBuf<double> &gmPrec<double>(phase_hip_solver *s) {
  return s->gmZ;
}
template <>
DBuf<float> &gmPrec<float>(phase_hip_solver *s) {
  return s->gmZ32;
}
template <class T>
KrylovVecs<T> &krylov(phase_hip_solver *s);
template <>
KrylovVecs<double> &krylov<double>(phase_hip_solver *s) {
  return s->k64;
}
template <>
KrylovVecs<float> &krylov<float>(phase_hip_solver *s) {
  return s->k32;
}
template <class T>
T *factor(phase_hip_solver *s);
template <>
double *factor<double>(phase_hip_solver *s) {
  return s->dFac64.p;
}
template <>
float *factor<float>(phase_hip_solver *s) {
  return s->dFac32.p;
}

// GMRES cycles with Krylov vectors in T from the current x until the FP64
// true residual meets the tolerance, `budget` steps are spent or (T = float)
// the cycles stall; returns the steps taken.  *stalled tells the last apart.
template <class T>
int gmresCycles(phase_hip_solver *s, double b2, int budget, bool *stalled) {
  constexpr bool kMixed = std::is_same<T, float>::value;
  const int n = s->rank, m = s->gmresRestart, ld = m + 1;
  hipStream_t st = s->stream;
  const IluDev<T> L = iluDev<T>(s, factor<T>(s));
  const int G = gridFor(n, kRedBlock, kRedCap);
  const int groups = (m + 1 + kGmGroup - 1) / kGmGroup;
  gmBasis<T>(s).alloc((size_t)(m + 1) * n);
  gmPrec<T>(s).alloc((size_t)m * n);
  s->gmH.alloc((size_t)ld * m);
  // small: cs (m), sn (m), g (m + 1), y (m), hcur (m + 1), vscale (1)
  s->gmSmall.alloc((size_t)5 * (m + 1) + 1);
  s->gmPart.alloc((size_t)groups * kGmGroup * G + G);
  krylov<T>(s).alloc(n);
  s->dR64.alloc(n);
  T *V = gmBasis<T>(s).p, *Z = gmPrec<T>(s).p;
  double *H = s->gmH.p;
  double *cs = s->gmSmall.p, *sn = cs + (m + 1), *g = sn + (m + 1), *y = g + (m + 1), *hcur = y + (m + 1);
  double *vscale = hcur + (m + 1);
  double *dotPart = s->gmPart.p, *normPart = dotPart + (size_t)groups * kGmGroup * G;
  Scalars *sc = s->dSc.p;
  const double tol2 = s->tol * s->tol * b2;
  const int ge = gridFor(n, 256, kElemCap);
  int total = 0, stall = 0;
  double bestRR = HUGE_VAL, startRR = -1., estRR = -1.;
  *stalled = false;
  while (true) {
    // FP64 true residual; it decides convergence
    hipLaunchKernelGGL((kResidual<double>), dim3(spmvGrid(s)), dim3(kRedBlock), 0, st, n, s->ord.nSlices,
                       s->dSliceOff.p, s->dSellCol.p, s->dSell64.p, s->dX.p, s->dB.p, s->dR64.p,
                       (double *)nullptr, s->dPartials.p);
    hipLaunchKernelGGL(kFinishResidual, dim3(1), dim3(kFinBlock), 0, st, sc, s->dPartials.p, spmvGrid(s));
    pullScalars(s);
    const double rr = s->hSc[0].rr;
    s->error = std::sqrt(rr / b2);
    ++s->outerPasses;
    if (!(rr > tol2) || total >= budget) break;
    if (kMixed) {
      // FP32 trouble: no new best, or a cycle that did not halve the true
      // residual while its own estimate claimed ten times more
      if (startRR > 0.) {
        const bool lost = rr > 0.25 * startRR && rr > 100. * estRR;
        stall = lost || !(rr < bestRR) ? stall + 1 : 0;
      }
      if (rr < bestRR) {
        bestRR = rr;
        PH_CHECK(hipMemcpyAsync(s->dXBest.p, s->dX.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
      }
      if (stall >= kGmMixedStall) {
        *stalled = true;
        break;
      }
      startRR = rr;
    }
    // the cycle's target: the tolerance, or (mixed) the FP32 cycle's drop
    hipLaunchKernelGGL(kGmSetTol, dim3(1), dim3(64), 0, st,
                       kMixed ? std::max(tol2, kGmMixedDrop * kGmMixedDrop * rr) : tol2, sc);
    hipLaunchKernelGGL(kGmStart, dim3(1), dim3(64), 0, st, m, g, vscale, sc);
    hipLaunchKernelGGL((kGmFirst<T>), dim3(ge), dim3(256), 0, st, n, (const double *)s->dR64.p,
                       (const double *)vscale, V);
    int j = 0, k = 0;
    while (j < m && total + j < budget) {
      const int steps = std::min({s->chunk, m - j, budget - total - j});
      for (int q = 0; q < steps; ++q, ++j) {
        T *vj = V + (size_t)j * n, *zj = Z + (size_t)j * n, *w = V + (size_t)(j + 1) * n;
        launchLU<T, kProNone>(s, L, sc, vj, krylov<T>(s).y.p, (T *)nullptr, (const T *)nullptr,
                              (const T *)nullptr, zj);
        if (s->precond == 1) mlFinishApply<T>(s, sc, vj, zj);
        hipLaunchKernelGGL((kSpmv<T, double>), dim3(spmvGrid(s)), dim3(kRedBlock), 0, st, n, s->ord.nSlices,
                           s->dSliceOff.p, s->dSellCol.p, s->dSell64.p, (const T *)zj, w);
        for (int pass = 0; pass < 2; ++pass) {
          hipLaunchKernelGGL((kGmDots<T>), dim3(G, (j + 1 + kGmGroup - 1) / kGmGroup), dim3(kRedBlock), 0, st, n,
                             j + 1, (const T *)V, (const T *)w, dotPart, (const Scalars *)sc);
          hipLaunchKernelGGL(kGmDotsFinish, dim3(1), dim3(kFinBlock), 0, st, j + 1, G, (const double *)dotPart,
                             hcur, H + (size_t)j * ld, pass, (const Scalars *)sc);
          if (pass == 0)
            hipLaunchKernelGGL((kGmAxpy<T, false>), dim3(G), dim3(kRedBlock), 0, st, n, j + 1, (const T *)V,
                               (const double *)hcur, w, normPart, (const Scalars *)sc);
          else
            hipLaunchKernelGGL((kGmAxpy<T, true>), dim3(G), dim3(kRedBlock), 0, st, n, j + 1, (const T *)V,
                               (const double *)hcur, w, normPart, (const Scalars *)sc);
        }
        hipLaunchKernelGGL(kGmGivens, dim3(1), dim3(kFinBlock), 0, st, j, ld, (const double *)normPart, G, H, cs,
                           sn, g, vscale, sc);
        hipLaunchKernelGGL((kGmScale<T>), dim3(ge), dim3(256), 0, st, n, w, (const double *)vscale,
                           (const Scalars *)sc);
      }
      PH_CHECK(hipGetLastError());
      pullScalars(s);
      k = s->hSc[0].iters;
      estRR = s->hSc[0].rr;
      if (s->hSc[0].done != kDoneRunning) break;
    }
    if (s->hSc[0].done == kDoneBreakdown) {
      // the column of the step that broke down is not finite: keep the steps
      // before it; with none (FP32: a factor or vector out of FP32 range) the
      // solve ends here, or continues in FP64
      k -= 1;
      if (k <= 0) {
        *stalled = kMixed;
        break;
      }
    }
    // x += Z y over the k steps of this cycle, then restart from the true residual
    hipLaunchKernelGGL(kGmSolveY, dim3(1), dim3(64), 0, st, k, ld, (const double *)H, (const double *)g, y);
    hipLaunchKernelGGL((kGmUpdateX<T>), dim3(G), dim3(kRedBlock), 0, st, n, k, (const T *)Z, (const double *)y,
                       s->dX.p);
    PH_CHECK(hipGetLastError());
    total += k;
  }
  if (*stalled && bestRR < HUGE_VAL) {
    PH_CHECK(hipMemcpyAsync(s->dX.p, s->dXBest.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    s->error = std::sqrt(bestRR / b2);
  }
  return total;
}

}  // namespace

void solveGmres(phase_hip_solver *s, double b2) {
  int total = 0;
  bool stalled = false;
  if (s->mode != PHASE_HIP_MODE_FP64) {
    ensureSingle(s);
    s->dXBest.alloc(s->rank);
    total = gmresCycles<float>(s, b2, s->maxIters, &stalled);
    if (stalled) ++s->fallbacks;
  }
  if (s->mode == PHASE_HIP_MODE_FP64 || stalled) {
    ensureFactor64(s);
    total += gmresCycles<double>(s, b2, s->maxIters - total, &stalled);
  }
  checkDeviceErr(s);
  s->innerIters = total;
  s->nIters = total;
}

}  // namespace phase
