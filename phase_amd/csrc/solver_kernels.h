// Device kernels of the sparse solver (included by solver.hip only).
//
// Data layout (all in "solver order", see ordering.cpp):
//  * A in SELL-64: slice s = rows [64s, 64s+64); entry e of lane l at
//    sliceOff[s] + 64*e + l; column -1 = padding.  Rows keep the caller's
//    in-row order, so y = A x sums each row in the same order as a plain CSR
//    loop over the caller's slots.
//  * ILU(0) factor fac = [L slots | U slots | diagonal]; block b (<= 64 rows,
//    one wavefront) keeps its L / U entries at lOff / uOff + 64*e + lane,
//    ascending column per row.
//  * Krylov vectors are dense arrays in solver order.
#pragma once

#include <hip/hip_runtime.h>
#include <type_traits>

#include "phase_internal.h"

namespace phase {

// The sweep kernels hold a row's factor entries in registers, W of each kind
// (in-block L, in-block U, off-block U; off-block L beyond W take a tail
// loop).  W = 4 covers triangle and quad meshes and halves the registers, so
// more blocks are resident per launch; W = 8 is the general case.
constexpr int kRedBlock = 256;

// The dispatcher hands workgroups to the 8 XCDs round robin (workgroup w runs
// on XCD w % 8), and each XCD has its own L2.  xcdTile(w, G) gives XCD k one
// contiguous range of the grid's G tiles instead of every eighth tile: under
// the bandwidth-reducing row order neighbouring tiles gather the same x lines,
// which then come from one L2 instead of up to eight.  A bijection of [0, G).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcdTile(int w, int G) {
  const int per = G / kXcds, rem = G % kXcds, x = w % kXcds;
  return x * per + min(x, rem) + w / kXcds;
}

// Prologues a forward sweep can form its right-hand side with (and store):
// BiCGStab's p and s, or CG's x and r update.
enum { kProNone = 0, kProP = 1, kProS = 2, kProCgR = 3 };

template <class T>
struct IluDev {
  const BlockInfo *blocks;
  const RowInfo *rows;
  const LaneRec *lanes;
  const int *lCol, *uCol;
  T *fac;
  int nb, nLSlots, nUSlots, n;
};

__device__ __forceinline__ double waveSum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Deterministic grid reduction: per-block fixed-order sums go to partials; the
// last block to finish (atomic ticket) sums the partials in index order.
// Returns true in the last block; thread 0 then holds the totals in v.
template <int NV>
__device__ bool reduceLast(double (&v)[NV], double *partials, unsigned *counter, int slot) {
  __shared__ double sred[NV][kRedBlock / 64];
  __shared__ bool sLast;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = waveSum(v[q]);
    if (lane == 0) sred[q][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.;
      for (int w = 0; w < kRedBlock / 64; ++w) s += sred[q][w];
      partials[q * gridDim.x + slot] = s;
    }
    __threadfence();
    unsigned t = atomicAdd(counter, 1u);
    sLast = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!sLast) return false;
  __threadfence();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = 0.;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += kRedBlock) s += partials[q * gridDim.x + i];
    s = waveSum(s);
    __syncthreads();
    if (lane == 0) sred[q][wave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double tsum = 0.;
      for (int w = 0; w < kRedBlock / 64; ++w) tsum += sred[q][w];
      v[q] = tsum;
    }
  }
  if (threadIdx.x == 0) *counter = 0u;
  return true;
}

// Fence-free deterministic reduction, part 1: each block writes its
// fixed-order sums to partials[q * gridDim.x + slot] (slot: the block's tile).  A one-block
// finisher kernel (launched next on the stream, so the kernel boundary orders
// the writes) adds them up in index order.  Unlike reduceLast this needs no
// device-scope fence per block.
template <int NV>
__device__ void blockPartials(double (&v)[NV], double *partials, int slot) {
  __shared__ double sred[NV][kRedBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = waveSum(v[q]);
    if (lane == 0) sred[q][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.;
      for (int w = 0; w < kRedBlock / 64; ++w) s += sred[q][w];
      partials[q * gridDim.x + slot] = s;
    }
  }
}

// One-block finishers run 1024 threads: a few thousand partials then take a
// single round of loads.
constexpr int kFinBlock = 1024;

// Part 2 (one block of BS threads): totals of nb partial blocks in index
// order; thread 0 holds them in v.
template <int NV, int BS = kRedBlock>
__device__ void finishPartials(const double *partials, int nb, double (&v)[NV]) {
  __shared__ double sred[NV][BS / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // eight loads per vector in flight per thread: the sum is bound by a few
  // memory round trips, not by one per partial
  constexpr int U = 8;
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.;
  for (int i0 = threadIdx.x; i0 < nb; i0 += U * BS) {
    double t[NV][U];
#pragma unroll
    for (int q = 0; q < NV; ++q)
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = i0 + u * BS;
        t[q][u] = i < nb ? partials[q * nb + i] : 0.;
      }
#pragma unroll
    for (int q = 0; q < NV; ++q)
#pragma unroll
      for (int u = 0; u < U; ++u) acc[q] += t[q][u];
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double s = waveSum(acc[q]);
    if (lane == 0) sred[q][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double t = 0.;
      for (int w = 0; w < BS / 64; ++w) t += sred[q][w];
      v[q] = t;
    }
  }
}

// ---------------------------------------------------------------- data movement
template <class T>
__global__ void kGather(int m, const int *__restrict__ map, const double *__restrict__ src,
                        T *__restrict__ dst) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    const int s = map[i];
    dst[i] = s >= 0 ? (T)src[s] : (T)0;
  }
}

// out[k] = in[perm[k]] (original -> solver order)
template <class TI, class TO>
__global__ void kPermuteIn(int n, const int *__restrict__ perm, const TI *__restrict__ in,
                           TO *__restrict__ out, double scale) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    out[k] = (TO)(scale * (double)in[perm[k]]);
}

// out[perm[k]] = in[k] (solver -> original order)
template <class T, class TO = double>
__global__ void kPermuteOut(int n, const int *__restrict__ perm, const T *__restrict__ in,
                            TO *__restrict__ out) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    out[perm[k]] = (TO)in[k];
}

template <class TI, class TO>
__global__ void kConvert(int n, const TI *__restrict__ in, TO *__restrict__ out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = (TO)in[i];
}

// ---------------------------------------------------------------- SpMV (SELL-64)
// Row `slice*64+lane` of A x.  The matrix values may be wider than the vectors
// (TV = double with T = float: FP64 coefficients, FP32 Krylov vectors); the
// products then accumulate in the wider type and round once at the end.
// The slice is uniform over the wavefront, so its width K is a scalar and the
// loads are issued as three batches -- columns and values, then the gathered
// x, then the fma chain in slot order -- rather than one dependent pair of
// round trips per slot (padding slots, column -1, read x[0] and add nothing).
constexpr int kSellBatch = 8;
template <class T, class TV>
__device__ __forceinline__ T sellRow(int slice, int lane, const int *__restrict__ sliceOff,
                                     const int *__restrict__ col, const TV *__restrict__ val,
                                     const T *__restrict__ x) {
  using A = typename std::conditional<(sizeof(TV) > sizeof(T)), TV, T>::type;
  slice = __builtin_amdgcn_readfirstlane(slice);
  const int o = sliceOff[slice];
  const int K = (sliceOff[slice + 1] - o) >> 6;
  A acc = 0;
  for (int e0 = 0; e0 < K; e0 += kSellBatch) {
    int c[kSellBatch];
    TV a[kSellBatch];
    // the matrix streams through once per product: non-temporal loads keep it
    // from evicting the gathered x lines from L2
#pragma unroll
    for (int q = 0; q < kSellBatch; ++q)
      if (e0 + q < K) {
        const int idx = o + ((e0 + q) << 6) + lane;
        c[q] = __builtin_nontemporal_load(col + idx);
        a[q] = __builtin_nontemporal_load(val + idx);
      }
    T xv[kSellBatch];
#pragma unroll
    for (int q = 0; q < kSellBatch; ++q)
      if (e0 + q < K) xv[q] = x[c[q] >= 0 ? c[q] : 0];
#pragma unroll
    for (int q = 0; q < kSellBatch; ++q)
      if (e0 + q < K && c[q] >= 0) acc = fma((A)a[q], (A)xv[q], acc);
  }
  return (T)acc;
}

// y = A x (matrix values in TV; rows accumulate in the wider of T and TV)
template <class T, class TV = T>
__global__ __launch_bounds__(kRedBlock) void kSpmv(int n, int nSlices, const int *__restrict__ sliceOff,
                                                   const int *__restrict__ col, const TV *__restrict__ val,
                                                   const T *__restrict__ x, T *__restrict__ y) {
  const int tile = xcdTile(blockIdx.x, gridDim.x);
  const int slice = tile * (kRedBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (slice >= nSlices) return;
  const int row = slice * 64 + lane;
  if (row >= n) return;
  y[row] = sellRow(slice, lane, sliceOff, col, val, x);
}

// v = A ph ; partials of (r0, v)   [kFinishV: alpha = rho / (r0, v)]
template <class T, class TV = T>
__global__ __launch_bounds__(kRedBlock) void kSpmvDotV(int n, int nSlices, const int *__restrict__ sliceOff,
                                                       const int *__restrict__ col, const TV *__restrict__ val,
                                                       const T *__restrict__ ph, T *__restrict__ v,
                                                       const T *__restrict__ r0, const Scalars *sc,
                                                       double *partials) {
  if (sc->done) return;
  const int tile = xcdTile(blockIdx.x, gridDim.x);
  const int slice = tile * (kRedBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int row = slice * 64 + lane;
  double d[1] = {0.};
  if (slice < nSlices && row < n) {
    const T y = sellRow(slice, lane, sliceOff, col, val, ph);
    v[row] = y;
    d[0] = (double)r0[row] * (double)y;
  }
  blockPartials<1>(d, partials, tile);
}

static __global__ __launch_bounds__(kFinBlock) void kFinishV(Scalars *sc, const double *partials, int nb) {
  if (sc->done) return;
  double d[1];
  finishPartials<1, kFinBlock>(partials, nb, d);
  if (threadIdx.x == 0) {
    sc->rv = d[0];
    if (d[0] == 0. || !isfinite(d[0])) sc->done = kDoneBreakdown;
    else sc->alpha = sc->rho / d[0];
  }
}

// t = A sh ; partials of (t, s), (t, t)   [kFinishT: omega = (t, s) / (t, t)]
template <class T, class TV = T>
__global__ __launch_bounds__(kRedBlock) void kSpmvDotT(int n, int nSlices, const int *__restrict__ sliceOff,
                                                       const int *__restrict__ col, const TV *__restrict__ val,
                                                       const T *__restrict__ sh, T *__restrict__ t,
                                                       const T *__restrict__ s, const Scalars *sc,
                                                       double *partials) {
  if (sc->done) return;
  const int tile = xcdTile(blockIdx.x, gridDim.x);
  const int slice = tile * (kRedBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int row = slice * 64 + lane;
  double d[2] = {0., 0.};
  if (slice < nSlices && row < n) {
    const T y = sellRow(slice, lane, sliceOff, col, val, sh);
    t[row] = y;
    d[0] = (double)y * (double)s[row];
    d[1] = (double)y * (double)y;
  }
  blockPartials<2>(d, partials, tile);
}

static __global__ __launch_bounds__(kFinBlock) void kFinishT(Scalars *sc, const double *partials, int nb) {
  if (sc->done) return;
  double d[2];
  finishPartials<2, kFinBlock>(partials, nb, d);
  if (threadIdx.x == 0) {
    sc->ts = d[0];
    sc->tt = d[1];
    sc->omega = d[1] > 0. ? d[0] / d[1] : 0.;
  }
}

// r = b - A x (optionally r0 = r); sc->rr = (r, r)
template <class T>
__global__ __launch_bounds__(kRedBlock) void kResidual(int n, int nSlices, const int *__restrict__ sliceOff,
                                                       const int *__restrict__ col, const T *__restrict__ val,
                                                       const T *__restrict__ x, const T *__restrict__ b,
                                                       T *__restrict__ r, T *__restrict__ r0, double *partials) {
  const int tile = xcdTile(blockIdx.x, gridDim.x);
  const int slice = tile * (kRedBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int row = slice * 64 + lane;
  double d[1] = {0.};
  if (slice < nSlices && row < n) {
    const T y = sellRow(slice, lane, sliceOff, col, val, x);
    const T ri = b[row] - y;
    r[row] = ri;
    if (r0) r0[row] = ri;
    d[0] = (double)ri * (double)ri;
  }
  blockPartials<1>(d, partials, tile);
}

// rr = (r, r) of the residual above
static __global__ __launch_bounds__(kFinBlock) void kFinishResidual(Scalars *sc, const double *partials, int nb) {
  double d[1];
  finishPartials<1, kFinBlock>(partials, nb, d);
  if (threadIdx.x == 0) sc->rr = d[0];
}

// *out = (v, v)
template <class T>
__global__ __launch_bounds__(kRedBlock) void kNorm2(int n, const T *__restrict__ v, double *partials) {
  double d[1] = {0.};
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock)
    d[0] += (double)v[i] * (double)v[i];
  blockPartials<1>(d, partials, blockIdx.x);
}

static __global__ __launch_bounds__(kFinBlock) void kFinishNorm2(const double *partials, int nb, double *out) {
  double d[1];
  finishPartials<1, kFinBlock>(partials, nb, d);
  if (threadIdx.x == 0) *out = d[0];
}

// ---------------------------------------------------------------- BiCGStab vector steps
template <class T>
__global__ void kBicgInit(int n, Scalars *sc, T *p, T *v, T *x, double tol2, int maxIters,
                          double minRatio) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    p[i] = 0;
    v[i] = 0;
    if (x) x[i] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->rho = sc->rr;
    sc->rrPrev = sc->rr;
    sc->alpha = 1.;
    sc->omega = 1.;
    sc->beta = 0.;
    sc->tol2 = tol2;
    sc->iters = 0;
    sc->maxIters = maxIters;
    sc->minRatio = minRatio;
    sc->done = sc->rr <= tol2 ? kDoneConverged : kDoneRunning;
  }
}

// x += alpha ph + omega sh ; r = s - omega t ; partials of (r0, r), (r, r).
// The solution x may be kept in a wider type (TX) than the Krylov vectors.
template <class T, class TX>
__global__ __launch_bounds__(kRedBlock) void kUpdate(int n, const Scalars *sc, TX *__restrict__ x,
                                                     const T *__restrict__ ph, const T *__restrict__ sh,
                                                     T *__restrict__ r, const T *__restrict__ s,
                                                     const T *__restrict__ t, const T *__restrict__ r0,
                                                     double *partials) {
  if (sc->done) return;
  const double alpha = sc->alpha, omega = sc->omega;
  if (!isfinite(alpha) || !isfinite(omega)) return;  // kFinishUpdate flags the breakdown
  const T om = (T)omega;
  const TX alx = (TX)alpha, omx = (TX)omega;
  double d[2] = {0., 0.};
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock) {
    const TX inc = alx * (TX)ph[i] + omx * (TX)sh[i];
    if (isfinite((double)inc)) x[i] = x[i] + inc;
    else d[1] = NAN;  // poisoned direction: reported as a breakdown, x untouched
    const T ri = s[i] - om * t[i];
    r[i] = ri;
    d[0] += (double)r0[i] * (double)ri;
    d[1] += (double)ri * (double)ri;
  }
  blockPartials<2>(d, partials, blockIdx.x);
}

// rho' = (r0, r), rr = (r, r), beta, stopping tests
static __global__ __launch_bounds__(kRedBlock) void kFinishUpdate(Scalars *sc, const double *partials, int nb) {
  if (sc->done) return;
  double d[2];
  finishPartials<2>(partials, nb, d);
  if (threadIdx.x != 0) return;
  if (!isfinite(sc->alpha) || !isfinite(sc->omega)) {
    sc->done = kDoneBreakdown;
    return;
  }
  const double rhoNew = d[0], rr = d[1];
  sc->rrPrev = sc->rr;
  sc->rr = rr;
  sc->beta = (rhoNew / sc->rho) * (sc->alpha / sc->omega);
  sc->rho = rhoNew;
  sc->iters += 1;
  if (!isfinite(rr) || !isfinite(rhoNew)) sc->done = kDoneBreakdown;
  else if (rr <= sc->tol2) sc->done = kDoneConverged;
  else if (rhoNew == 0. || sc->omega == 0.) sc->done = kDoneBreakdown;
  else if (sc->minRatio > 0. && rr > sc->minRatio * sc->minRatio * sc->rrPrev) sc->done = kDoneStagnation;
  else if (sc->iters >= sc->maxIters) sc->done = kDoneMaxIters;
}

// ---------------------------------------------------------------- CG vector steps
// Preconditioned conjugate gradients (Hestenes-Stiefel, with M = ILU(0)) for
// symmetric systems.  One iteration: q = A p with (p, q) and (r, r)
// (kSpmvDotCg, kFinishCgQ: stopping test, alpha); x += alpha p and
// r -= alpha q inside the forward sweeps that apply M^{-1} to the new r
// (kProCgR); (r, z) partials (kDotRZ); beta and p = z + beta p (kCgDirectionBeta).

// q = A p ; partials of (p, q) and of (r, r) for the stopping test (r as the
// previous iteration's sweeps left it)
template <class T, class TV = T>
__global__ __launch_bounds__(kRedBlock) void kSpmvDotCg(int n, int nSlices, const int *__restrict__ sliceOff,
                                                        const int *__restrict__ col, const TV *__restrict__ val,
                                                        const T *__restrict__ p, T *__restrict__ q,
                                                        const T *__restrict__ r, const Scalars *sc,
                                                        double *partials) {
  if (sc->done) return;
  const int tile = xcdTile(blockIdx.x, gridDim.x);
  const int slice = tile * (kRedBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int row = slice * 64 + lane;
  double d[2] = {0., 0.};
  if (slice < nSlices && row < n) {
    const T y = sellRow(slice, lane, sliceOff, col, val, p);
    q[row] = y;
    const double ri = (double)r[row];
    d[0] = (double)p[row] * (double)y;
    d[1] = ri * ri;
  }
  blockPartials<2>(d, partials, tile);
}

// Stopping tests on rr = (r, r) (the residual of the current x), then
// alpha = rho / (p, q); the sweeps that follow move x and r by alpha.
static __global__ __launch_bounds__(kFinBlock) void kFinishCgQ(Scalars *sc, const double *partials, int nb) {
  if (sc->done) return;
  double d[2];
  finishPartials<2, kFinBlock>(partials, nb, d);
  if (threadIdx.x != 0) return;
  const double pq = d[0], rr = d[1];
  sc->rho = sc->rhoNext;
  sc->rrPrev = sc->rr;
  sc->rr = rr;
  sc->rv = pq;
  if (!isfinite(rr)) sc->done = kDoneBreakdown;
  else if (rr <= sc->tol2) sc->done = kDoneConverged;
  else if (sc->iters > 0 && sc->minRatio > 0. && rr > sc->minRatio * sc->minRatio * sc->rrPrev)
    sc->done = kDoneStagnation;
  else if (sc->iters >= sc->maxIters) sc->done = kDoneMaxIters;
  else if (pq == 0. || !isfinite(pq)) sc->done = kDoneBreakdown;
  else {
    sc->alpha = sc->rho / pq;
    sc->iters += 1;
  }
}

// partials of (r, z), and with FLEX also of (z, v): v still holds A p of the
// previous direction
template <class T, bool FLEX = false>
__global__ __launch_bounds__(kRedBlock) void kDotRZ(int n, const Scalars *sc, const T *__restrict__ r,
                                                    const T *__restrict__ z, const T *__restrict__ v,
                                                    double *partials) {
  if (sc->done) return;
  double d[FLEX ? 2 : 1];
  d[0] = 0.;
  if (FLEX) d[FLEX ? 1 : 0] = 0.;
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock) {
    d[0] += (double)r[i] * (double)z[i];
    if (FLEX) d[FLEX ? 1 : 0] += (double)z[i] * (double)v[i];
  }
  blockPartials<FLEX ? 2 : 1>(d, partials, blockIdx.x);
}

// rho' = (r, z), beta and p = z + beta p in one launch: every block adds up
// the partials itself, in the same fixed order, so all blocks hold the same
// beta; block 0 records beta and rho' (as rhoNext: rho itself is still being
// read by the other blocks; the next alpha takes it over).  At the start p
// and v are zero.  Plain CG: beta = rho' / rho.  Flexible CG (a preconditioner
// that varies from apply to apply, Notay 2000): beta = -(z, A p) / (p, A p)
// with (p, A p) from the last alpha (sc->rv), which keeps p A-orthogonal to the
// previous direction.
template <class T, bool FLEX = false>
__global__ __launch_bounds__(kRedBlock) void kCgDirectionBeta(int n, Scalars *sc, const double *partials, int nb,
                                                              const T *__restrict__ z, T *__restrict__ p) {
  if (sc->done) return;
  double d[FLEX ? 2 : 1];
  finishPartials<FLEX ? 2 : 1>(partials, nb, d);
  __shared__ double sBeta;
  __shared__ int sBad;
  if (threadIdx.x == 0) {
    const double rhoNew = d[0], rho = sc->rho;
    double beta;
    bool bad;
    if (FLEX) {
      const double zq = d[FLEX ? 1 : 0], pq = sc->rv;
      beta = zq == 0. ? 0. : -zq / pq;
      bad = !isfinite(rhoNew) || rhoNew == 0. || !isfinite(beta);
    } else {
      bad = !isfinite(rhoNew) || rhoNew == 0. || rho == 0.;
      beta = bad ? 0. : rhoNew / rho;
    }
    sBad = bad;
    sBeta = bad ? 0. : beta;
    if (blockIdx.x == 0) {
      if (bad) {
        sc->done = kDoneBreakdown;
      } else {
        sc->beta = sBeta;
        sc->rhoNext = rhoNew;
      }
    }
  }
  __syncthreads();
  if (sBad) return;
  const T beta = (T)sBeta;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = z[i] + beta * p[i];
}

// r32 = r64 (reliable update of the recursive residual); inner->rr = outer->rr
template <class T>
__global__ void kReplaceResidual(int n, const double *__restrict__ r64, T *__restrict__ r, Scalars *inner,
                                 const Scalars *outer) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) r[i] = (T)r64[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) inner->rr = outer->rr;
}

template <class TI, class TO>
__global__ void kConvert2(int n, const TI *__restrict__ in, TO *__restrict__ a, TO *__restrict__ b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const TO v = (TO)in[i];
    a[i] = v;
    b[i] = v;
  }
}

// q = A e ; sc->ts = (r, q), sc->tt = (q, q)
template <class T>
__global__ __launch_bounds__(kRedBlock) void kSpmvDot2(int n, int nSlices, const int *__restrict__ sliceOff,
                                                       const int *__restrict__ col, const T *__restrict__ val,
                                                       const T *__restrict__ e, T *__restrict__ q,
                                                       const T *__restrict__ r, Scalars *sc, double *partials,
                                                       unsigned *counter) {
  const int tile = xcdTile(blockIdx.x, gridDim.x);
  const int slice = tile * (kRedBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int row = slice * 64 + lane;
  double d[2] = {0., 0.};
  if (slice < nSlices && row < n) {
    const T y = sellRow(slice, lane, sliceOff, col, val, e);
    q[row] = y;
    d[0] = (double)r[row] * (double)y;
    d[1] = (double)y * (double)y;
  }
  if (reduceLast<2>(d, partials, counter, tile) && threadIdx.x == 0) {
    sc->ts = d[0];
    sc->tt = d[1];
  }
}

// alpha = ts / tt ; x += alpha e ; r -= alpha q ; outer->rr = (r, r)
template <class T>
__global__ __launch_bounds__(kRedBlock) void kMinResUpdate(int n, const Scalars *step, Scalars *outer,
                                                           T *__restrict__ x, const T *__restrict__ e,
                                                           T *__restrict__ r, const T *__restrict__ q,
                                                           double *partials, unsigned *counter) {
  const T alpha = (T)(step->ts / step->tt);
  double d[1] = {0.};
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock) {
    x[i] += alpha * e[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    d[0] += (double)ri * (double)ri;
  }
  if (reduceLast<1>(d, partials, counter, blockIdx.x) && threadIdx.x == 0) outer->rr = d[0];
}

// x64 += scale * x32 with scale = sqrt(*rrOuter) (undoing the inner normalisation)
template <class T>
__global__ void kAccumulate(int n, double *__restrict__ x, const T *__restrict__ e, const Scalars *outer) {
  const double sc = sqrt(outer->rr);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    x[i] += sc * (double)e[i];
}

// r32 = r64 / ||r64|| ; r0 = r32
template <class T>
__global__ void kNormalise(int n, const double *__restrict__ r64, T *__restrict__ r, T *__restrict__ r0,
                           const Scalars *outer) {
  const double inv = 1.0 / sqrt(outer->rr);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const T v = (T)(r64[i] * inv);
    r[i] = v;
    r0[i] = v;
  }
}

// ---------------------------------------------------------------- ILU(0)
// A sweep or factorisation launch covers the blocks of one colour: blocks of
// one colour share no coupling, so block = blockBase + xcdTile(blockIdx.x)
// needs no waiting -- kernel boundaries order the colours.
__device__ __forceinline__ int colourBlock(int blockBase) { return blockBase + xcdTile(blockIdx.x, gridDim.x); }

// Factorisation, IKJ order per row (the oracle's orc_ilu0): for each L entry
// (i,k) in ascending k: l = a_ik / u_kk, then a_ij -= l * u_kj over U(k) in
// ascending j.  Rows of earlier blocks are final (earlier colour launch); rows
// of the same block are final after their level.  A zero pivot sets bit 2 of
// *err.
template <class T>
__global__ __launch_bounds__(64) void kIluFactor(IluDev<T> L, const int *__restrict__ updPtr,
                                                 const int *__restrict__ updSrc,
                                                 const int *__restrict__ updDst, int *err, int blockBase) {
  const int b = colourBlock(blockBase);
  const BlockInfo B = L.blocks[b];
  const int lane = threadIdx.x;
  const bool act = lane < B.nRows;
  const int row = B.rowStart + lane;
  const RowInfo ri = act ? L.rows[row] : RowInfo{};
  T *fac = L.fac;
  const int dbase = L.nLSlots + L.nUSlots;
  for (int lev = 0; lev < B.nLev; ++lev) {
    if (act && ri.lev == lev) {
      for (int e = 0; e < ri.nL; ++e) {
        const int s = B.lOff + (e << 6) + lane;
        const int k = L.lCol[s];
        const T lik = fac[s] / fac[dbase + k];
        fac[s] = lik;
        for (int u = updPtr[s]; u < updPtr[s + 1]; ++u) {
          const int dst = updDst[u];
          fac[dst] = fma(-lik, fac[updSrc[u]], fac[dst]);
        }
      }
      if (fac[dbase + row] == (T)0) atomicOr(err, 2);
    }
    __syncthreads();
  }
}

// A sweep launch is latency-bound: every block runs one wavefront and a
// colour's blocks are all resident at once, so the launch lasts about as long
// as one wavefront's chain of dependent memory round trips.  The kernels keep
// that chain at three: (1) block record, lane record and the run's scalars;
// (2) the row's own vector entries and every factor value and column index it
// needs; (3) the solution values of earlier-coloured rows.  The in-block
// triangular solve then runs level by level through LDS.

// fac[dbase + n + i] = 1 / u_ii (read by the backward sweep)
template <class T>
__global__ void kInvDiag(int n, T *fac, int dbase) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    fac[dbase + n + i] = (T)1 / fac[dbase + i];
}

// Forward sweep body for one block: returns y of the lane's row.  PRO selects
// the fused Krylov prologue that forms the right-hand side on the fly:
// kProP: p = r + beta (p - omega v);  kProS: s = r - alpha v (both stored);
// kProCgR: r = r - alpha q with r and q in pv and v, and x += alpha p with p
// in `in` (r and x stored).
// Accumulation order (earlier columns first) is that of the oracle's forward
// substitution, so the result is bit-identical to it.
template <class T, int PRO, int W>
__device__ __forceinline__ T lBody(const IluDev<T> &L, const BlockInfo &B, bool act, int row, unsigned info, int lane, const Scalars *sc,
                                   const T *__restrict__ in, const T *__restrict__ y, const T *__restrict__ pv,
                                   const T *__restrict__ r, const T *__restrict__ v, T *ylds, T &pro,
                                   const double *__restrict__ xs, double &xo) {
  const int nIn = act ? (int)(info & 15u) : 0;
  const int nOff = act ? (int)((info >> 4) & 63u) : 0;
  const int lev = (int)((info >> 18) & 63u);
  // the row's own entries load unconditionally (idle lanes read row 0 and
  // discard it): a branch here would make the compiler wait for them before
  // the factor loads below
  const int rs = act ? row : 0;
  T acc;
  if (PRO == kProP) {
    const T beta = (T)sc->beta, omega = (T)sc->omega;
    acc = r[rs] + beta * (pv[rs] - omega * v[rs]);
  } else if (PRO == kProS) {
    const T alpha = (T)sc->alpha;
    acc = r[rs] - alpha * v[rs];
  } else if (PRO == kProCgR) {
    const double alpha = sc->alpha;
    acc = pv[rs] - (T)alpha * v[rs];
    xo = xs[rs] + alpha * (double)in[rs];
  } else {
    acc = in[rs];
  }
  pro = acc;  // stored by the caller at the end: a store here would wait for the loads
  if (!act) acc = 0;
  T lv[W];
  int lc[W];
#pragma unroll
  for (int q = 0; q < W; ++q) {
    lv[q] = 0;
    lc[q] = 0;
    if (q < nIn) {
      const int s = B.lOff + ((nOff + q) << 6) + lane;
      lv[q] = L.fac[s];
      lc[q] = L.lCol[s];
    }
  }
  // column indices stay 32-bit unsigned until the gathers: a sign extension
  // here would make each predicated load wait for its own result
  T ov[W];
  unsigned oc[W];
#pragma unroll
  for (int q = 0; q < W; ++q) {
    ov[q] = 0;
    oc[q] = 0;
    if (q < nOff) {
      const int s = B.lOff + (q << 6) + lane;
      ov[q] = L.fac[s];
      oc[q] = (unsigned)L.lCol[s];
    }
  }
  // keep every load above in flight before the gathers below are addressed
  // (otherwise each column load is waited for in turn)
  asm volatile("" ::: "memory");
  T oy[W];
#pragma unroll
  for (int q = 0; q < W; ++q) oy[q] = q < nOff ? y[oc[q]] : (T)0;
#pragma unroll
  for (int q = 0; q < W; ++q)
    if (q < nOff) acc = fma(-ov[q], oy[q], acc);
  for (int e = W; e < nOff; ++e) {
    const int s = B.lOff + (e << 6) + lane;
    acc = fma(-L.fac[s], y[L.lCol[s]], acc);
  }
  for (int l = 0; l < B.nLev; ++l) {
    if (act && lev == l) {
#pragma unroll
      for (int q = 0; q < W; ++q)
        if (q < nIn) acc = fma(-lv[q], ylds[lc[q] - B.rowStart], acc);
      if constexpr (W == 8) {
        // rows with more in-block entries than registers (coarse multigrid levels)
        for (int e = W; e < nIn; ++e) {
          const int s = B.lOff + ((nOff + e) << 6) + lane;
          acc = fma(-L.fac[s], ylds[L.lCol[s] - B.rowStart], acc);
        }
      }
      ylds[lane] = acc;
    }
    __syncthreads();
  }
  return acc;
}

// Backward sweep of one block in two parts: uLoad issues every load the row
// needs from the factor, uSolve gathers later-coloured solution values and
// runs the levels from acc = y of the lane's row, returning x of the row.
// In-block columns precede later-block columns, as in the oracle's back
// substitution.
template <class T, int W>
struct UPre {
  T d;  // 1 / u_ii
  int nIn, nOff, ulev;
  T uv[W];
  int uc[W];
  T ou[W];
  unsigned oc[W];
};

template <class T, int W>
__device__ __forceinline__ void uLoad(const IluDev<T> &L, const BlockInfo &B, bool act, int row, unsigned info,
                                      int lane, UPre<T, W> &u) {
  const T *facU = L.fac + L.nLSlots;
  u.nIn = act ? (int)((info >> 10) & 15u) : 0;
  u.nOff = act ? (int)((info >> 14) & 15u) : 0;
  u.ulev = (int)((info >> 24) & 63u);
  // inverse diagonal, stored after the diagonal: unconditional load, as in lBody
  const T d0 = L.fac[L.nLSlots + L.nUSlots + L.n + (act ? row : 0)];
  u.d = act ? d0 : (T)1;
#pragma unroll
  for (int q = 0; q < W; ++q) {
    u.uv[q] = 0;
    u.uc[q] = 0;
    if (q < u.nIn) {
      const int s = B.uOff + (q << 6) + lane;
      u.uv[q] = facU[s];
      u.uc[q] = L.uCol[s];
    }
  }
#pragma unroll
  for (int q = 0; q < W; ++q) {
    u.ou[q] = 0;
    u.oc[q] = 0;
    if (q < u.nOff) {
      const int s = B.uOff + ((u.nIn + q) << 6) + lane;
      u.ou[q] = facU[s];
      u.oc[q] = (unsigned)L.uCol[s];
    }
  }
}

template <class T, int W>
__device__ __forceinline__ T uSolve(const IluDev<T> &L, const BlockInfo &B, bool act, int lane, T acc,
                                    const T *__restrict__ x, T *xlds, const UPre<T, W> &u) {
  asm volatile("" ::: "memory");
  T ox[W];
#pragma unroll
  for (int q = 0; q < W; ++q) ox[q] = q < u.nOff ? x[u.oc[q]] : (T)0;
  // later-block terms are final before the levels start: their sum leaves the
  // level loop, whose body (issued for the whole wave at every level) keeps
  // only the in-block terms and one multiply by the inverse diagonal
  T soff = 0;
#pragma unroll
  for (int q = 0; q < W; ++q)
    if (q < u.nOff) soff = fma(u.ou[q], ox[q], soff);
  const T *facU = L.fac + L.nLSlots;
  if constexpr (W == 8) {
    // rows with more off-block entries than registers (coarse multigrid levels)
    for (int e = W; e < u.nOff; ++e) {
      const int s = B.uOff + ((u.nIn + e) << 6) + lane;
      soff = fma(facU[s], x[L.uCol[s]], soff);
    }
  }
  for (int l = 0; l < B.nULev; ++l) {
    if (act && u.ulev == l) {
#pragma unroll
      for (int q = 0; q < W; ++q)
        if (q < u.nIn) acc = fma(-u.uv[q], xlds[u.uc[q] - B.rowStart], acc);
      if constexpr (W == 8) {
        for (int e = W; e < u.nIn; ++e) {
          const int s = B.uOff + (e << 6) + lane;
          acc = fma(-facU[s], xlds[L.uCol[s] - B.rowStart], acc);
        }
      }
      acc = (acc - soff) * u.d;
      xlds[lane] = acc;
    }
    __syncthreads();
  }
  return acc;
}

// Forward sweep y = L^{-1} b (unit L).  After convergence (sc->done) every
// lane is idle and the launch only walks the block's levels.
template <class T, int PRO, int W = 8>
__global__ __launch_bounds__(64) void kLSolve(IluDev<T> L, int blockBase, const Scalars *sc,
                                              const T *__restrict__ in, T *__restrict__ y, T *__restrict__ pv,
                                              const T *__restrict__ r, const T *__restrict__ v,
                                              double *__restrict__ xs) {
  __shared__ T ylds[64];
  const int b = colourBlock(blockBase);
  const int lane = threadIdx.x;
  const int done = sc->done;
  const BlockInfo B = L.blocks[b];
  const LaneRec rec = L.lanes[(size_t)b * 64 + lane];
  const bool act = rec.row >= 0 && !done;
  T pro;
  double xo = 0.;
  const T acc = lBody<T, PRO, W>(L, B, act, rec.row, rec.info, lane, sc, in, y, pv, r, v, ylds, pro, xs, xo);
  if (act) {
    if (PRO != kProNone) pv[rec.row] = pro;
    if (PRO == kProCgR) xs[rec.row] = xo;
    y[rec.row] = acc;
  }
}

// Backward sweep x = U^{-1} y (colours in reverse order).  With addTo set,
// the finished x of each row is also added to addFrom's value into addTo
// (z = z + M^-1 t of a multigrid post-smoothing, without a launch of its own;
// addFrom may be addTo).
template <class T, int W = 8>
__global__ __launch_bounds__(64) void kUSolve(IluDev<T> L, int blockBase, const Scalars *sc,
                                              const T *__restrict__ y, T *__restrict__ x,
                                              const T *addFrom, T *addTo) {
  __shared__ T xlds[64];
  const int b = colourBlock(blockBase);
  const int lane = threadIdx.x;
  const int done = sc->done;
  const BlockInfo B = L.blocks[b];
  const LaneRec rec = L.lanes[(size_t)b * 64 + lane];
  const bool act = rec.row >= 0 && !done;
  const T y0 = y[act ? rec.row : 0];
  UPre<T, W> u;
  uLoad<T, W>(L, B, act, rec.row, rec.info, lane, u);
  const T acc = uSolve<T, W>(L, B, act, lane, act ? y0 : (T)0, x, xlds, u);
  if (act) {
    x[rec.row] = acc;
    if (addTo) addTo[rec.row] = addFrom[rec.row] + acc;
  }
}

// Last colour of a preconditioner apply: blocks of
// the last colour have no later-coloured neighbours, so the forward and the
// backward sweep of such a block run back to back in one launch.  The U
// entries load after the forward levels: issued up front they overlap the
// forward sweep's loads but double the registers, and the lower occupancy
// costs more (a second round of blocks per launch) than the round trip.
template <class T, int PRO, int W = 8>
__global__ __launch_bounds__(64) void kLUSolve(IluDev<T> L, int blockBase, const Scalars *sc,
                                               const T *__restrict__ in, const T *__restrict__ y,
                                               T *__restrict__ pv, const T *__restrict__ r,
                                               const T *__restrict__ v, T *__restrict__ x, double *__restrict__ xs,
                                               const T *addFrom, T *addTo) {
  __shared__ T ylds[64];
  __shared__ T xlds[64];
  const int b = colourBlock(blockBase);
  const int lane = threadIdx.x;
  const int done = sc->done;
  const BlockInfo B = L.blocks[b];
  const LaneRec rec = L.lanes[(size_t)b * 64 + lane];
  const bool act = rec.row >= 0 && !done;
  T pro;
  double xo = 0.;
  const T yv = lBody<T, PRO, W>(L, B, act, rec.row, rec.info, lane, sc, in, y, pv, r, v, ylds, pro, xs, xo);
  UPre<T, W> u;
  uLoad<T, W>(L, B, act, rec.row, rec.info, lane, u);
  const T acc = uSolve<T, W>(L, B, act, lane, yv, x, xlds, u);
  if (act) {
    if (PRO != kProNone) pv[rec.row] = pro;
    if (PRO == kProCgR) xs[rec.row] = xo;
    x[rec.row] = acc;
    if (addTo) addTo[rec.row] = addFrom[rec.row] + acc;
  }
}

}  // namespace phase
