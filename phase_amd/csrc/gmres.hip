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
// FP64 throughout.  The Hessenberg column, the rotations and the residual
// estimate live on the device, so an Arnoldi step needs no host round trip;
// the host checks the step count and the convergence flag after every chunk
// of steps, and every launch after convergence returns at once (except the
// SpMV).  A cycle that ends on the estimate is followed by the true residual,
// which decides.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "multilevel.h"
#include "solver.h"
#include "solver_kernels.h"
#include "solver_launch.h"

namespace phase {

namespace {

constexpr int kGmGroup = 8;

// partials of (w, V_j) for the kGmGroup vectors j = 8 blockIdx.y + q < k
__global__ __launch_bounds__(kRedBlock) void kGmDots(int n, int k, const double *__restrict__ V,
                                                     const double *__restrict__ w, double *partials,
                                                     const Scalars *sc) {
  if (sc->done) return;
  const int g0 = blockIdx.y * kGmGroup;
  double d[kGmGroup];
#pragma unroll
  for (int q = 0; q < kGmGroup; ++q) d[q] = 0.;
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock) {
    const double wi = w[i];
#pragma unroll
    for (int q = 0; q < kGmGroup; ++q)
      if (g0 + q < k) d[q] = fma(V[(size_t)(g0 + q) * n + i], wi, d[q]);
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

// w -= sum_{j<k} h_j V_j; with NORM the partials of (w, w)
template <bool NORM>
__global__ __launch_bounds__(kRedBlock) void kGmAxpy(int n, int k, const double *__restrict__ V,
                                                     const double *__restrict__ hcur, double *__restrict__ w,
                                                     double *partials, const Scalars *sc) {
  if (sc->done) return;
  __shared__ double h[kGmMaxRestart + 1];
  for (int j = threadIdx.x; j < k; j += kRedBlock) h[j] = hcur[j];
  __syncthreads();
  double d[1] = {0.};
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock) {
    double v = w[i];
    for (int j = 0; j < k; ++j) v = fma(-h[j], V[(size_t)j * n + i], v);
    w[i] = v;
    if (NORM) d[0] = fma(v, v, d[0]);
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
__global__ void kGmScale(int n, double *__restrict__ v, const double *scale, const Scalars *sc) {
  if (sc->done) return;
  const double f = *scale;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] *= f;
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

// x += sum_{j<k} y_j Z_j
__global__ __launch_bounds__(kRedBlock) void kGmUpdateX(int n, int k, const double *__restrict__ Z,
                                                        const double *__restrict__ y, double *__restrict__ x) {
  __shared__ double sy[kGmMaxRestart];
  for (int j = threadIdx.x; j < k; j += kRedBlock) sy[j] = y[j];
  __syncthreads();
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock) {
    double v = x[i];
    for (int j = 0; j < k; ++j) v = fma(sy[j], Z[(size_t)j * n + i], v);
    x[i] = v;
  }
}

}  // namespace

void solveGmres(phase_hip_solver *s, double b2) {
  const int n = s->rank, m = s->gmresRestart, ld = m + 1;
  hipStream_t st = s->stream;
  ensureFactor64(s);
  const IluDev<double> L = iluDev<double>(s, s->dFac64.p);
  const int G = gridFor(n, kRedBlock, kRedCap);
  const int groups = (m + 1 + kGmGroup - 1) / kGmGroup;
  s->gmV.alloc((size_t)(m + 1) * n);
  s->gmZ.alloc((size_t)m * n);
  s->gmH.alloc((size_t)ld * m);
  // small: cs (m), sn (m), g (m + 1), y (m), hcur (m + 1), vscale (1)
  s->gmSmall.alloc((size_t)5 * (m + 1) + 1);
  s->gmPart.alloc((size_t)groups * kGmGroup * G + G);
  s->k64.alloc(n);
  double *V = s->gmV.p, *Z = s->gmZ.p, *H = s->gmH.p;
  double *cs = s->gmSmall.p, *sn = cs + (m + 1), *g = sn + (m + 1), *y = g + (m + 1), *hcur = y + (m + 1);
  double *vscale = hcur + (m + 1);
  double *dotPart = s->gmPart.p, *normPart = dotPart + (size_t)groups * kGmGroup * G;
  Scalars *sc = s->dSc.p;
  hipLaunchKernelGGL(kGmSetTol, dim3(1), dim3(64), 0, st, s->tol * s->tol * b2, sc);
  const int ge = gridFor(n, 256, kElemCap);
  int total = 0;
  while (true) {
    // true residual into v_0, then the cycle's start
    hipLaunchKernelGGL((kResidual<double>), dim3(spmvGrid(s)), dim3(kRedBlock), 0, st, n, s->ord.nSlices,
                       s->dSliceOff.p, s->dSellCol.p, s->dSell64.p, s->dX.p, s->dB.p, V, (double *)nullptr,
                       s->dPartials.p);
    hipLaunchKernelGGL(kFinishResidual, dim3(1), dim3(kFinBlock), 0, st, sc, s->dPartials.p, spmvGrid(s));
    hipLaunchKernelGGL(kGmStart, dim3(1), dim3(64), 0, st, m, g, vscale, sc);
    hipLaunchKernelGGL(kGmScale, dim3(ge), dim3(256), 0, st, n, V, (const double *)vscale, (const Scalars *)sc);
    pullScalars(s);
    s->error = std::sqrt(s->hSc[0].rr / b2);
    ++s->outerPasses;
    if (s->hSc[0].done != kDoneRunning || total >= s->maxIters) break;
    int j = 0, k = 0;
    while (j < m && total + j < s->maxIters) {
      const int steps = std::min({s->chunk, m - j, s->maxIters - total - j});
      for (int q = 0; q < steps; ++q, ++j) {
        double *vj = V + (size_t)j * n, *zj = Z + (size_t)j * n, *w = V + (size_t)(j + 1) * n;
        launchLU<double, kProNone>(s, L, sc, vj, s->k64.y.p, (double *)nullptr, (const double *)nullptr,
                                   (const double *)nullptr, zj);
        if (s->precond == 1) mlFinishApply<double>(s, sc, vj, zj);
        hipLaunchKernelGGL((kSpmv<double>), dim3(spmvGrid(s)), dim3(kRedBlock), 0, st, n, s->ord.nSlices,
                           s->dSliceOff.p, s->dSellCol.p, s->dSell64.p, (const double *)zj, w);
        for (int pass = 0; pass < 2; ++pass) {
          hipLaunchKernelGGL(kGmDots, dim3(G, (j + 1 + kGmGroup - 1) / kGmGroup), dim3(kRedBlock), 0, st, n,
                             j + 1, (const double *)V, (const double *)w, dotPart, (const Scalars *)sc);
          hipLaunchKernelGGL(kGmDotsFinish, dim3(1), dim3(kFinBlock), 0, st, j + 1, G, (const double *)dotPart,
                             hcur, H + (size_t)j * ld, pass, (const Scalars *)sc);
          if (pass == 0)
            hipLaunchKernelGGL((kGmAxpy<false>), dim3(G), dim3(kRedBlock), 0, st, n, j + 1, (const double *)V,
                               (const double *)hcur, w, normPart, (const Scalars *)sc);
          else
            hipLaunchKernelGGL((kGmAxpy<true>), dim3(G), dim3(kRedBlock), 0, st, n, j + 1, (const double *)V,
                               (const double *)hcur, w, normPart, (const Scalars *)sc);
        }
        hipLaunchKernelGGL(kGmGivens, dim3(1), dim3(kFinBlock), 0, st, j, ld, (const double *)normPart, G, H, cs,
                           sn, g, vscale, sc);
        hipLaunchKernelGGL(kGmScale, dim3(ge), dim3(256), 0, st, n, w, (const double *)vscale,
                           (const Scalars *)sc);
      }
      PH_CHECK(hipGetLastError());
      pullScalars(s);
      k = s->hSc[0].iters;
      if (s->hSc[0].done != kDoneRunning) break;
    }
    if (s->hSc[0].done == kDoneBreakdown && k == 0) break;
    // x += Z y over the k steps of this cycle, then restart from the true residual
    hipLaunchKernelGGL(kGmSolveY, dim3(1), dim3(64), 0, st, k, ld, (const double *)H, (const double *)g, y);
    hipLaunchKernelGGL(kGmUpdateX, dim3(G), dim3(kRedBlock), 0, st, n, k, (const double *)Z, (const double *)y,
                       s->dX.p);
    PH_CHECK(hipGetLastError());
    // the next pass's true residual decides convergence (kGmStart resets the flag)
    total += k;
  }
  checkDeviceErr(s);
  s->innerIters = total;
  s->nIters = total;
}

}  // namespace phase
