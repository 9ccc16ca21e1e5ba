// Aggregation multigrid preconditioner: hierarchy, numeric set-up and the
// cycle.  The mathematics is written out in multilevel.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "multilevel.h"
#include "solver.h"
#include "solver_kernels.h"
#include "solver_launch.h"

namespace phase {

namespace {

// Device scalars of the K cycle's two flexible CG iterations at one level.
struct MlScalars {
  double pq1, alpha1, gamma, alpha2;
};

template <class T>
struct MlVecs {
  // r: right-hand side, e: solution of this level's coarse solve; t, w, y, u:
  // cycle work; z1, q1, r1, z2: the K cycle's iterations
  DBuf<T> r, e, t, w, y, u, z1, q1, r1, z2;
  void alloc(int n) {
    for (DBuf<T> *b : {&r, &e, &t, &w, &y, &u, &z1, &q1, &r1, &z2}) b->alloc(n);
  }
  // the finest level uses the cycle work vectors only
  void allocCycle(int n) {
    for (DBuf<T> *b : {&t, &w, &y}) b->alloc(n);
  }
};

}  // namespace

struct MlLevel {
  int n = 0;
  phase_hip_solver *sol = nullptr;  // nested solver; null for the dense last level
  bool jacobi = false;              // damped Jacobi smoothing (dinv) instead of the ILU(0) factor
  DBuf<double> dinv64;              // omega / a_ii in solver order
  DBuf<float> dinv32;
  // maps from the finer level
  DBuf<int> galPtr, galSrc;         // coarse slot q = sum of finer caller slots galSrc[galPtr[q] ..]
  DBuf<int> restRow;                // this level's solver row -> aggregate of the finer level (dense: identity)
  DBuf<int> finerAggPtr;            // finer solver rows of each aggregate
  DBuf<int> pIdx;                   // finer solver row -> this level's solver row
  int nFiner = 0, nSlots = 0;
  // dense last level
  DBuf<int> cRowPtr, cColInd;
  DBuf<double> cVals, dense, inv;
  MlVecs<double> v64;
  MlVecs<float> v32;
  DBuf<MlScalars> msc;
  DBuf<double> partials, partials2;
};

struct Multilevel {
  std::vector<MlLevel> lv;  // lv[0]: first coarse level
  int jacobiFrom = 0;       // first Jacobi-smoothed level the hierarchy was built with
  bool valuesValid = false, singleValid = false;
  MlVecs<double> f64;  // fine-level work vectors (t, w, y used)
  MlVecs<float> f32;
  std::vector<int> sizes;
};

namespace {

template <class T>
MlVecs<T> &vecs(MlLevel &L);
template <>
MlVecs<double> &vecs<double>(MlLevel &L) {
  return L.v64;
}
template <>
MlVecs<float> &vecs<float>(MlLevel &L) {
  return L.v32;
}
template <class T>
MlVecs<T> &fineVecs(Multilevel &M);
template <>
MlVecs<double> &fineVecs<double>(Multilevel &M) {
  return M.f64;
}
template <>
MlVecs<float> &fineVecs<float>(Multilevel &M) {
  return M.f32;
}
template <class T>
const T *dinvOf(const MlLevel &L);
template <>
const double *dinvOf<double>(const MlLevel &L) {
  return L.dinv64.p;
}
template <>
const float *dinvOf<float>(const MlLevel &L) {
  return L.dinv32.p;
}
// the ILU(0) factor a level applies in precision T
template <class T>
T *facOf(phase_hip_solver *s);
template <>
double *facOf<double>(phase_hip_solver *s) {
  return s->dFac64.p;
}
template <>
float *facOf<float>(phase_hip_solver *s) {
  return s->dFac32.p;
}

// ---------------------------------------------------------------- kernels
// coarse[q] = sum of fine[src[ptr[q] .. ptr[q+1])] in list order
__global__ void kGalerkin(int m, const int *__restrict__ ptr, const int *__restrict__ src,
                          const double *__restrict__ fine, double *__restrict__ coarse) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < m; q += gridDim.x * blockDim.x) {
    double acc = 0.;
    for (int k = ptr[q]; k < ptr[q + 1]; ++k) acc += fine[src[k]];
    coarse[q] = acc;
  }
}

// dense row-major n x n from CSR (one thread per row)
__global__ void kDenseFill(int n, const int *__restrict__ rowPtr, const int *__restrict__ colInd,
                           const double *__restrict__ vals, double *__restrict__ a) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    for (int j = 0; j < n; ++j) a[(size_t)i * n + j] = 0.;
    for (int k = rowPtr[i]; k < rowPtr[i + 1]; ++k) a[(size_t)i * n + colInd[k]] += vals[k];
  }
}

// inv = a^-1 by Gauss-Jordan elimination with partial pivoting, one
// workgroup (a is destroyed).  A zero pivot sets bit 4 of *err.
constexpr int kDenseMax = 512;
constexpr int kGjBlock = 1024;
__global__ __launch_bounds__(kGjBlock) void kDenseInverse(int n, double *__restrict__ a, double *__restrict__ inv,
                                                          int *err) {
  __shared__ double fcol[kDenseMax], prow[kDenseMax], pinv[kDenseMax];
  __shared__ double sval[kGjBlock / 64];
  __shared__ int sidx[kGjBlock / 64];
  __shared__ int sPivot;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int t = tid; t < n * n; t += kGjBlock) inv[t] = (t / n == t % n) ? 1. : 0.;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    // pivot: largest |a[i][k]|, i >= k (lowest i on a tie)
    double best = -1.;
    int bi = k;
    for (int i = k + tid; i < n; i += kGjBlock) {
      const double v = fabs(a[(size_t)i * n + k]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ov > best || (ov == best && oi < bi)) {
        best = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      sval[wave] = best;
      sidx[wave] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      double b = sval[0];
      int p = sidx[0];
      for (int w = 1; w < kGjBlock / 64; ++w)
        if (sval[w] > b || (sval[w] == b && sidx[w] < p)) {
          b = sval[w];
          p = sidx[w];
        }
      if (!(b > 0.)) atomicOr(err, 4);
      sPivot = p;
    }
    __syncthreads();
    const int p = sPivot;
    // swap rows k and p, then scale row k by 1 / pivot (kept in LDS)
    for (int j = tid; j < n; j += kGjBlock) {
      const double ak = a[(size_t)k * n + j], ap = a[(size_t)p * n + j];
      const double ik = inv[(size_t)k * n + j], ip = inv[(size_t)p * n + j];
      prow[j] = ap;
      pinv[j] = ip;
      if (p != k) {
        a[(size_t)p * n + j] = ak;
        inv[(size_t)p * n + j] = ik;
      }
    }
    __syncthreads();
    const double d = prow[k];
    const double rd = d != 0. ? 1. / d : 0.;
    __syncthreads();  // every thread has the pivot before the row is scaled
    for (int j = tid; j < n; j += kGjBlock) {
      prow[j] *= rd;
      pinv[j] *= rd;
    }
    for (int i = tid; i < n; i += kGjBlock) fcol[i] = i == k ? 0. : a[(size_t)i * n + k];
    __syncthreads();
    for (int j = tid; j < n; j += kGjBlock) {
      a[(size_t)k * n + j] = prow[j];
      inv[(size_t)k * n + j] = pinv[j];
    }
    // eliminate column k from every other row
    for (int t = tid; t < n * n; t += kGjBlock) {
      const int i = t / n, j = t - i * n;
      const double f = fcol[i];
      if (f != 0.) {
        a[t] = fma(-f, prow[j], a[t]);
        inv[t] = fma(-f, pinv[j], inv[t]);
      }
    }
    __syncthreads();
  }
}

// The same inverse with the matrix in registers, for n <= 32 RB: the 1024
// threads form a 32 x 32 grid and thread (tr, tc) holds the entries (tr + 32 a,
// tc + 32 b), a, b < RB.  In-place Gauss-Jordan with partial pivoting and no
// row exchanges: step k takes as pivot the unused row p_k with the largest
// |a_ik| (lowest i on a tie) and eliminates column k from every other row,
// replacing it by the column of the inverse being built.  That is the
// unpivoted in-place inversion of B = P A (B row k = A row p_k), so physical
// row p_k ends holding row k of B^-1 = A^-1 P^T: inv[k][p_j] = R[p_k][j].
// Two barriers per step (column k and the pivot candidates published; pivot
// row published); the column, candidate and row buffers alternate between
// steps, so a step never overwrites what a slower thread of the previous step
// still reads.
template <int RB>
__global__ __launch_bounds__(kGjBlock) void kDenseInverseReg(int n, const double *__restrict__ a0,
                                                             double *__restrict__ inv, int *err) {
  constexpr int S = 32 * RB;
  // entries past n of the column and row buffers stay 0, so the update needs
  // no bounds tests: padding entries of A stay 0 or unchanged and are never stored
  __shared__ double fcol[2][S], prow[2][S], candV[2][32];
  __shared__ int candI[2][32], pivOf[S], kOf[S], used[S];
  const int tid = threadIdx.x, tr = tid >> 5, tc = tid & 31;
  double A[RB][RB];
#pragma unroll
  for (int a = 0; a < RB; ++a)
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const int i = tr + 32 * a, j = tc + 32 * b;
      A[a][b] = (i < n && j < n) ? a0[(size_t)i * n + j] : 0.;
    }
  for (int i = tid; i < S; i += kGjBlock) {
    used[i] = 0;
    pivOf[i] = 0;
    kOf[i] = 0;
    fcol[0][i] = fcol[1][i] = prow[0][i] = prow[1][i] = 0.;
  }
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    const int buf = k & 1, kb = k >> 5, kc = k & 31;
    // the 32 threads holding column k publish it, each with its best pivot
    // candidate among its unused rows (largest |a_ik|, lowest i on a tie; a
    // NaN is taken, so some unused row always is, and flagged below); b == kb
    // is the same for every thread, so only the branch taken is issued
#pragma unroll
    for (int b = 0; b < RB; ++b)
      if (b == kb && tc == kc) {
        double best = -1.;
        int bi = n;
#pragma unroll
        for (int a = 0; a < RB; ++a) {
          const int i = tr + 32 * a;
          if (i < n) {
            fcol[buf][i] = A[a][b];
            const double av = fabs(A[a][b]);
            if (!used[i] && !(av <= best)) {
              best = av;
              bi = i;
            }
          }
        }
        candV[buf][tr] = best;
        candI[buf][tr] = bi;
      }
    __syncthreads();
    // every thread reduces the 32 candidates in the same order
    double best = -1.;
    int bi = n;
#pragma unroll 8
    for (int q = 0; q < 32; ++q) {
      const double v = candV[buf][q];
      const int i = candI[buf][q];
      if (i < n && (!(v <= best) || (v == best && i < bi))) {
        best = v;
        bi = i;
      }
    }
    const int p = bi < n ? bi : 0;
    const double pv = fcol[buf][p];
    const double pivinv = pv != 0. ? 1. / pv : 0.;
    if (tid == 0) {
      // before the barrier: the next step's column owners read used[]
      used[p] = 1;
      pivOf[k] = p;
      kOf[p] = k;
      if (!(fabs(pv) > 0.)) atomicOr(err, 4);
    }
    // pivot row, scaled, with 1 / pivot in column k (row k of the inverse so far)
    const int pb = p >> 5, pr = p & 31;
#pragma unroll
    for (int a = 0; a < RB; ++a)
      if (a == pb && tr == pr) {
#pragma unroll
        for (int b = 0; b < RB; ++b) {
          const int j = tc + 32 * b;
          if (j < n) prow[buf][j] = A[a][b] * pivinv;
        }
        if (tc == kc) prow[buf][k] = pivinv;
      }
    __syncthreads();
    // column k becomes the inverse's column: a_ik = 0, then every row takes
    // a_ij -= a_ik_old * prow_j; the pivot row is then set to prow
#pragma unroll
    for (int b = 0; b < RB; ++b)
      if (b == kb && tc == kc) {
#pragma unroll
        for (int a = 0; a < RB; ++a) A[a][b] = 0.;
      }
    double rs[RB], f[RB];
#pragma unroll
    for (int b = 0; b < RB; ++b) rs[b] = prow[buf][tc + 32 * b];
#pragma unroll
    for (int a = 0; a < RB; ++a) f[a] = fcol[buf][tr + 32 * a];
#pragma unroll
    for (int a = 0; a < RB; ++a)
#pragma unroll
      for (int b = 0; b < RB; ++b) A[a][b] = fma(-f[a], rs[b], A[a][b]);
#pragma unroll
    for (int a = 0; a < RB; ++a)
      if (a == pb && tr == pr) {
#pragma unroll
        for (int b = 0; b < RB; ++b) A[a][b] = rs[b];
      }
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < RB; ++a)
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const int i = tr + 32 * a, j = tc + 32 * b;
      if (i < n && j < n) inv[(size_t)kOf[i] * n + pivOf[j]] = A[a][b];
    }
}

// The same register layout without pivoting: step k eliminates with the
// diagonal a_kk, so the column and the row of the step are known before any
// search and one barrier per step publishes both.  Gauss-Jordan without
// pivoting is stable for the diagonally dominant and the symmetric positive
// definite matrices the pressure equations coarsen to; a pivot below 1e-14 of
// the largest entry, or not finite, sets bit 8 of *err and the caller inverts
// again with pivoting.
template <int RB>
__global__ __launch_bounds__(kGjBlock) void kDenseInverseDiag(int n, const double *__restrict__ a0,
                                                              double *__restrict__ inv, int *err) {
  constexpr int S = 32 * RB;
  __shared__ double fcol[2][S], prow[2][S], smax[kGjBlock / 64];
  const int tid = threadIdx.x, tr = tid >> 5, tc = tid & 31, lane = tid & 63, wave = tid >> 6;
  double A[RB][RB];
  double m = 0.;
#pragma unroll
  for (int a = 0; a < RB; ++a)
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const int i = tr + 32 * a, j = tc + 32 * b;
      A[a][b] = (i < n && j < n) ? a0[(size_t)i * n + j] : 0.;
      m = fmax(m, fabs(A[a][b]));
    }
  for (int i = tid; i < S; i += kGjBlock) fcol[0][i] = fcol[1][i] = prow[0][i] = prow[1][i] = 0.;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if (lane == 0) smax[wave] = m;
  __syncthreads();
  double amax = 0.;
  for (int w = 0; w < kGjBlock / 64; ++w) amax = fmax(amax, smax[w]);
  const double tiny = 1e-14 * amax;
  for (int k = 0; k < n; ++k) {
    const int buf = k & 1, kb = k >> 5, kc = k & 31;
    // column k and row k as they stand (k >> 5 is the same for every thread:
    // only the branch taken is issued)
#pragma unroll
    for (int b = 0; b < RB; ++b)
      if (b == kb && tc == kc) {
#pragma unroll
        for (int a = 0; a < RB; ++a) {
          const int i = tr + 32 * a;
          if (i < n) fcol[buf][i] = A[a][b];
        }
      }
#pragma unroll
    for (int a = 0; a < RB; ++a)
      if (a == kb && tr == kc) {
#pragma unroll
        for (int b = 0; b < RB; ++b) {
          const int j = tc + 32 * b;
          if (j < n) prow[buf][j] = A[a][b];
        }
      }
    __syncthreads();
    const double pv = fcol[buf][k];
    const double pivinv = pv != 0. ? 1. / pv : 0.;
    if (tid == 0 && !(fabs(pv) > tiny)) atomicOr(err, 8);
    double rs[RB], f[RB];
#pragma unroll
    for (int b = 0; b < RB; ++b) rs[b] = (b == kb && tc == kc) ? pivinv : prow[buf][tc + 32 * b] * pivinv;
#pragma unroll
    for (int a = 0; a < RB; ++a) f[a] = fcol[buf][tr + 32 * a];
#pragma unroll
    for (int b = 0; b < RB; ++b)
      if (b == kb && tc == kc) {
#pragma unroll
        for (int a = 0; a < RB; ++a) A[a][b] = 0.;
      }
#pragma unroll
    for (int a = 0; a < RB; ++a)
#pragma unroll
      for (int b = 0; b < RB; ++b) A[a][b] = fma(-f[a], rs[b], A[a][b]);
#pragma unroll
    for (int a = 0; a < RB; ++a)
      if (a == kb && tr == kc) {
#pragma unroll
        for (int b = 0; b < RB; ++b) A[a][b] = rs[b];
      }
  }
#pragma unroll
  for (int a = 0; a < RB; ++a)
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const int i = tr + 32 * a, j = tc + 32 * b;
      if (i < n && j < n) inv[(size_t)i * n + j] = A[a][b];
    }
}

// Largest n the register Gauss-Jordan kernels take (RB = 6: 36 entries per thread).
constexpr int kDenseRegMax = 192;

// inv = dense^-1 into `inv`; err bit 4 on a singular matrix.  Up to
// kDenseRegMax rows, without pivoting first unless `pivot`, and with pivoting
// when that flags a small pivot (one host synchronisation each).
void denseInverse(hipStream_t st, int n, double *dense, double *inv, int *err, bool pivot) {
  auto readErr = [&] {
    int e = 0;
    PH_CHECK(hipMemcpyAsync(&e, err, sizeof(int), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipStreamSynchronize(st));
    return e;
  };
  if (n > kDenseRegMax) {
    hipLaunchKernelGGL(kDenseInverse, dim3(1), dim3(kGjBlock), 0, st, n, dense, inv, err);
    PH_CHECK(hipGetLastError());
    return;
  }
  if (!pivot) {
    if (n <= 64) hipLaunchKernelGGL(kDenseInverseDiag<2>, dim3(1), dim3(kGjBlock), 0, st, n, dense, inv, err);
    else if (n <= 128) hipLaunchKernelGGL(kDenseInverseDiag<4>, dim3(1), dim3(kGjBlock), 0, st, n, dense, inv, err);
    else hipLaunchKernelGGL(kDenseInverseDiag<6>, dim3(1), dim3(kGjBlock), 0, st, n, dense, inv, err);
    PH_CHECK(hipGetLastError());
    if (!(readErr() & 8)) return;
    PH_CHECK(hipMemsetAsync(err, 0, sizeof(int), st));
  }
  if (n <= 64) hipLaunchKernelGGL(kDenseInverseReg<2>, dim3(1), dim3(kGjBlock), 0, st, n, dense, inv, err);
  else if (n <= 128) hipLaunchKernelGGL(kDenseInverseReg<4>, dim3(1), dim3(kGjBlock), 0, st, n, dense, inv, err);
  else hipLaunchKernelGGL(kDenseInverseReg<6>, dim3(1), dim3(kGjBlock), 0, st, n, dense, inv, err);
  PH_CHECK(hipGetLastError());
}

// e = inv r (one wavefront per row, FP64 products)
template <class T>
__global__ __launch_bounds__(256) void kDenseApply(int n, const double *__restrict__ inv, const T *__restrict__ r,
                                                   T *__restrict__ e, const Scalars *sc) {
  if (sc->done) return;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= n) return;
  double acc = 0.;
  for (int j = lane; j < n; j += 64) acc = fma(inv[(size_t)row * n + j], (double)r[j], acc);
  acc = waveSum(acc);
  if (lane == 0) e[row] = (T)acc;
}

// t = b - A x (SELL-64, one row per lane)
template <class T, class TV>
__global__ __launch_bounds__(kRedBlock) void kMlResid(int n, int nSlices, const int *__restrict__ sliceOff,
                                                      const int *__restrict__ col, const TV *__restrict__ val,
                                                      const T *__restrict__ x, const T *__restrict__ b,
                                                      T *__restrict__ t, const Scalars *sc) {
  if (sc->done) return;
  const int tile = xcdTile(blockIdx.x, gridDim.x);
  const int slice = tile * (kRedBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int row = slice * 64 + lane;
  if (slice >= nSlices || row >= n) return;
  t[row] = b[row] - sellRow<T, TV>(slice, lane, sliceOff, col, val, x);
}

// rc[k] = sum of t over the finer rows of aggregate restRow[k] (identity when
// null; an aggregate's rows are consecutive in the finer solver order), in
// FP64: 16 lanes per coarse row, lane l summing rows l, l + 16, ... in order,
// then a fixed xor tree over the 16 lanes (one or two load rounds for an
// aggregate of about 32 rows, where a thread per aggregate took seven).
template <class T>
__global__ __launch_bounds__(256) void kMlRestrictRows(int nc, const int *__restrict__ restRow,
                                                       const int *__restrict__ aggPtr, const T *__restrict__ t,
                                                       T *__restrict__ rc, const Scalars *sc) {
  if (sc->done) return;
  const int k = blockIdx.x * 16 + (threadIdx.x >> 4), l = threadIdx.x & 15;
  double d = 0.;
  if (k < nc) {
    const int a = restRow ? restRow[k] : k;
    const int r1 = aggPtr[a + 1];
    for (int i = aggPtr[a] + l; i < r1; i += 16) d += (double)t[i];
  }
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) d += __shfl_xor(d, o, 16);
  if (k < nc && l == 0) rc[k] = (T)d;
}

// e[i] += ec[pIdx[i]]
template <class T>
__global__ void kMlProlongAdd(int n, const int *__restrict__ pIdx, const T *__restrict__ ec, T *__restrict__ e,
                              const Scalars *sc) {
  if (sc->done) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) e[i] += ec[pIdx[i]];
}

// Row `row` of A x in SELL-64 for a lane whose slice need not be the rest of
// the wavefront's: the products and their order are sellRow's (x[c] comes
// from gx(c)), so the result is bit-identical to it.
template <class T, class TV, class GX>
__device__ __forceinline__ T sellRowAt(int row, const int *__restrict__ sliceOff, const int *__restrict__ col,
                                       const TV *__restrict__ val, GX &&gx) {
  using A = typename std::conditional<(sizeof(TV) > sizeof(T)), TV, T>::type;
  const int slice = row >> 6, lane = row & 63;
  const int o = sliceOff[slice];
  const int K = (sliceOff[slice + 1] - o) >> 6;
  A acc = 0;
  for (int e0 = 0; e0 < K; e0 += 4) {
    int c[4];
    TV a[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (e0 + q < K) {
        const int idx = o + ((e0 + q) << 6) + lane;
        c[q] = col[idx];
        a[q] = val[idx];
      }
    T xv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (e0 + q < K) xv[q] = gx(c[q] >= 0 ? c[q] : 0);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (e0 + q < K && c[q] >= 0) acc = fma((A)a[q], (A)xv[q], acc);
  }
  return (T)acc;
}

// Coarse levels (launch-bound), kMlResid and the restriction in one launch:
// rc[k] = sum over the rows of aggregate restRow[k] of (b - A x), one
// wavefront per coarse row; each row's residual is rounded to T before the
// FP64 sum (in a wavefront's fixed order).
template <class T, class TV>
__global__ __launch_bounds__(256) void kMlResidRestrict(int nc, const int *__restrict__ restRow,
                                                        const int *__restrict__ aggPtr,
                                                        const int *__restrict__ sliceOff,
                                                        const int *__restrict__ col, const TV *__restrict__ val,
                                                        const T *__restrict__ x, const T *__restrict__ b,
                                                        T *__restrict__ rc, const Scalars *sc) {
  if (sc->done) return;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (k >= nc) return;
  const int a = restRow ? restRow[k] : k;
  const int row = aggPtr[a] + lane;
  double d = 0.;
  if (row < aggPtr[a + 1]) {
    const T t = b[row] - sellRowAt<T, TV>(row, sliceOff, col, val, [&](int c) { return x[c]; });
    d = (double)t;
  }
  d = waveSum(d);
  if (lane == 0) rc[k] = (T)d;
}

// Coarse levels, kMlProlongAdd and kMlResid in one launch:
// u = z + P ec and t = b - A u, where A's products read z + P ec at the
// gathered columns (the sums kMlProlongAdd would store) and z stays as it is.
template <class T, class TV>
__global__ __launch_bounds__(kRedBlock) void kMlResidProlong(int n, int nSlices, const int *__restrict__ sliceOff,
                                                             const int *__restrict__ col, const TV *__restrict__ val,
                                                             const int *__restrict__ pIdx,
                                                             const T *__restrict__ ec, const T *__restrict__ z,
                                                             const T *__restrict__ b, T *__restrict__ t,
                                                             T *__restrict__ u, const Scalars *sc) {
  if (sc->done) return;
  const int slice = blockIdx.x * (kRedBlock / 64) + (threadIdx.x >> 6);
  const int row = slice * 64 + (threadIdx.x & 63);
  if (slice >= nSlices || row >= n) return;
  const T y = sellRowAt<T, TV>(row, sliceOff, col, val, [&](int c) { return (T)(z[c] + ec[pIdx[c]]); });
  t[row] = b[row] - y;
  u[row] = z[row] + ec[pIdx[row]];
}

// q = A p ; partials of (p, q) and (p, r)
template <class T, class TV>
__global__ __launch_bounds__(kRedBlock) void kMlSpmvDot2(int n, int nSlices, const int *__restrict__ sliceOff,
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
    const T y = sellRow<T, TV>(slice, lane, sliceOff, col, val, p);
    q[row] = y;
    d[0] = (double)p[row] * (double)y;
    d[1] = (double)p[row] * (double)r[row];
  }
  blockPartials<2>(d, partials, tile);
}

// alpha = (p, r) / (p, q) from the partials, summed in the same fixed order
// in every block (every block holds the same alpha)
__device__ __forceinline__ double mlAlpha(const double *partials, int nb, double &pq) {
  double d[2];
  finishPartials<2>(partials, nb, d);
  __shared__ double sA, sPq;
  if (threadIdx.x == 0) {
    sPq = d[0];
    sA = (d[0] != 0. && isfinite(d[0]) && isfinite(d[1])) ? d[1] / d[0] : 0.;
  }
  __syncthreads();
  pq = sPq;
  return sA;
}

// first iteration: alpha1; x = alpha1 z1 ; r1 = r0 - alpha1 q1 (block 0 keeps (z1, q1))
template <class T>
__global__ __launch_bounds__(kRedBlock) void kMlFirst(int n, const Scalars *sc, MlScalars *m,
                                                      const double *partials, int nb, const T *__restrict__ z1,
                                                      const T *__restrict__ q1, const T *__restrict__ r0,
                                                      T *__restrict__ x, T *__restrict__ r1) {
  if (sc->done) return;
  double pq;
  const T a = (T)mlAlpha(partials, nb, pq);
  if (blockIdx.x == 0 && threadIdx.x == 0) m->pq1 = pq;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] = a * z1[i];
    r1[i] = r0[i] - a * q1[i];
  }
}

// partials of (z2, q1)
template <class T>
__global__ __launch_bounds__(kRedBlock) void kMlDotZQ(int n, const Scalars *sc, const T *__restrict__ z2,
                                                      const T *__restrict__ q1, double *partials) {
  if (sc->done) return;
  double d[1] = {0.};
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock)
    d[0] += (double)z2[i] * (double)q1[i];
  blockPartials<1>(d, partials, blockIdx.x);
}

// gamma = (z2, q1) / (z1, q1) ; z2 -= gamma z1 in every block (the same fixed-order sum)
template <class T>
__global__ __launch_bounds__(kRedBlock) void kMlDirection(int n, const Scalars *sc, const MlScalars *m,
                                                          const double *partials, int nb, const T *__restrict__ z1,
                                                          T *__restrict__ z2) {
  if (sc->done) return;
  double d[1];
  finishPartials<1>(partials, nb, d);
  __shared__ double sGamma;
  if (threadIdx.x == 0) {
    const double pq = m->pq1;
    sGamma = (pq != 0. && isfinite(d[0])) ? d[0] / pq : 0.;
  }
  __syncthreads();
  const T g = (T)sGamma;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) z2[i] -= g * z1[i];
}

// second iteration: alpha2 ; x += alpha2 p2
template <class T>
__global__ __launch_bounds__(kRedBlock) void kMlSecond(int n, const Scalars *sc, const double *partials, int nb,
                                                       const T *__restrict__ p2, T *__restrict__ x) {
  if (sc->done) return;
  double pq;
  const T a = (T)mlAlpha(partials, nb, pq);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] += a * p2[i];
}

// dinv[i] = omega / a_ii (the diagonal found in the row's SELL-64 entries)
__global__ void kMlDiagInv(int n, const int *__restrict__ sliceOff, const int *__restrict__ col,
                           const double *__restrict__ val, double omega, double *__restrict__ dinv) {
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < n; row += gridDim.x * blockDim.x) {
    const int slice = row >> 6, lane = row & 63;
    const int o = sliceOff[slice], K = (sliceOff[slice + 1] - o) >> 6;
    double d = 0.;
    for (int e = 0; e < K; ++e)
      if (col[o + (e << 6) + lane] == row) d += val[o + (e << 6) + lane];
    dinv[row] = d != 0. ? omega / d : 0.;
  }
}

// Jacobi pre-smoothing, residual and restriction in one launch (one
// wavefront per coarse row): z = dinv r on the aggregate's rows (stored),
// rc[k] = sum over them of r - A z, with A's products reading dinv_j r_j.
template <class T, class TV>
__global__ __launch_bounds__(256) void kMlJacResidRestrict(int nc, const int *__restrict__ restRow,
                                                           const int *__restrict__ aggPtr,
                                                           const int *__restrict__ sliceOff,
                                                           const int *__restrict__ col, const TV *__restrict__ val,
                                                           const T *__restrict__ dinv, const T *__restrict__ r,
                                                           T *__restrict__ z, T *__restrict__ rc,
                                                           const Scalars *sc) {
  if (sc->done) return;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (k >= nc) return;
  const int a = restRow ? restRow[k] : k;
  const int row = aggPtr[a] + lane;
  double d = 0.;
  if (row < aggPtr[a + 1]) {
    z[row] = dinv[row] * r[row];
    const T t = r[row] - sellRowAt<T, TV>(row, sliceOff, col, val, [&](int c) { return dinv[c] * r[c]; });
    d = (double)t;
  }
  d = waveSum(d);
  if (lane == 0) rc[k] = (T)d;
}

// Prolongation, residual and Jacobi post-smoothing in one launch:
// u = z + P ec; out = u + dinv (r - A u) (z and out distinct).  With DOT, also
// the per-block partials of (out, q) (the K cycle's (z2, q1)).
template <class T, class TV, bool DOT>
__global__ __launch_bounds__(kRedBlock) void kMlJacResidProlong(int n, int nSlices, const int *__restrict__ sliceOff,
                                                                const int *__restrict__ col,
                                                                const TV *__restrict__ val,
                                                                const int *__restrict__ pIdx,
                                                                const T *__restrict__ ec, const T *__restrict__ z,
                                                                const T *__restrict__ r, const T *__restrict__ dinv,
                                                                T *__restrict__ out, const Scalars *sc,
                                                                const T *__restrict__ q, double *partials) {
  if (sc->done) return;
  const int slice = blockIdx.x * (kRedBlock / 64) + (threadIdx.x >> 6);
  const int row = slice * 64 + (threadIdx.x & 63);
  double d[1] = {0.};
  if (slice < nSlices && row < n) {
    const T y = sellRowAt<T, TV>(row, sliceOff, col, val, [&](int c) { return (T)(z[c] + ec[pIdx[c]]); });
    const T u = z[row] + ec[pIdx[row]];
    const T o = u + dinv[row] * (r[row] - y);
    out[row] = o;
    if (DOT) d[0] = (double)o * (double)q[row];
  }
  if (DOT) blockPartials<1>(d, partials, blockIdx.x);
}

// The K cycle on a Jacobi level, second apply, first launch: kMlFirst's work
// (alpha1 = (z1, r) / (z1, q1) from kMlSpmvDot2's partials, summed in every
// block; e = alpha1 z1; r1 = r - alpha1 q1, stored) and kMlJacResidRestrict on
// r1, whose products form r1 at the gathered columns.  The same arithmetic as
// the two kernels.
template <class T, class TV>
__global__ __launch_bounds__(256) void kMlJacFirstRestrict(int nc, const int *__restrict__ restRow,
                                                           const int *__restrict__ aggPtr,
                                                           const int *__restrict__ sliceOff,
                                                           const int *__restrict__ col, const TV *__restrict__ val,
                                                           const T *__restrict__ dinv, const T *__restrict__ r,
                                                           const T *__restrict__ q1, const T *__restrict__ z1,
                                                           const double *partials, int nb, MlScalars *m,
                                                           T *__restrict__ r1, T *__restrict__ e, T *__restrict__ z,
                                                           T *__restrict__ rc, const Scalars *sc) {
  if (sc->done) return;
  double pq;
  const T a = (T)mlAlpha(partials, nb, pq);
  if (blockIdx.x == 0 && threadIdx.x == 0) m->pq1 = pq;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (k >= nc) return;
  const int ag = restRow ? restRow[k] : k;
  const int row = aggPtr[ag] + lane;
  double d = 0.;
  if (row < aggPtr[ag + 1]) {
    const T rr = r[row] - a * q1[row];
    r1[row] = rr;
    e[row] = a * z1[row];
    z[row] = dinv[row] * rr;
    const T t = rr - sellRowAt<T, TV>(row, sliceOff, col, val, [&](int c) { return dinv[c] * (r[c] - a * q1[c]); });
    d = (double)t;
  }
  d = waveSum(d);
  if (lane == 0) rc[k] = (T)d;
}

// The K cycle's direction on a Jacobi level: gamma = (z2, q1) / (z1, q1) (the
// partials of the second apply, summed in every block), p2 = z2 - gamma z1
// (stored), and the partials of (p2, A p2) and (p2, r1), A p2 formed from p2
// at the gathered columns (kMlDirection and kMlSpmvDot2 in one launch).
template <class T, class TV>
__global__ __launch_bounds__(kRedBlock) void kMlJacDirection(int n, int nSlices, const int *__restrict__ sliceOff,
                                                             const int *__restrict__ col, const TV *__restrict__ val,
                                                             const Scalars *sc, const MlScalars *m,
                                                             const double *partialsIn, int nb,
                                                             const T *__restrict__ z1, const T *__restrict__ z2,
                                                             const T *__restrict__ r1, T *__restrict__ p2,
                                                             double *partialsOut) {
  if (sc->done) return;
  double dz[1];
  finishPartials<1>(partialsIn, nb, dz);
  __shared__ double sGamma;
  if (threadIdx.x == 0) {
    const double pq = m->pq1;
    sGamma = (pq != 0. && isfinite(dz[0])) ? dz[0] / pq : 0.;
  }
  __syncthreads();
  const T g = (T)sGamma;
  const int slice = blockIdx.x * (kRedBlock / 64) + (threadIdx.x >> 6);
  const int row = slice * 64 + (threadIdx.x & 63);
  double d[2] = {0., 0.};
  if (slice < nSlices && row < n) {
    const T p = z2[row] - g * z1[row];
    p2[row] = p;
    const T y = sellRowAt<T, TV>(row, sliceOff, col, val, [&](int c) { return (T)(z2[c] - g * z1[c]); });
    d[0] = (double)p * (double)y;
    d[1] = (double)p * (double)r1[row];
  }
  blockPartials<2>(d, partialsOut, blockIdx.x);
}

// ---------------------------------------------------------------- host side
phase_hip_solver *makeNested(phase_hip_solver *s) {
  auto *c = new phase_hip_solver();
  c->device = s->device;
  c->stream = s->stream;  // borrowed: every level runs on the parent's stream
  c->blockRows = s->mlCoarseRows;
  c->geometric = s->geometric;
  return c;
}

void freeNested(phase_hip_solver *c) {
  if (!c) return;
  c->stream = nullptr;  // not ours to destroy
  phase_hip_solver_destroy(c);
}

int elemGrid(long n) { return gridFor(n, 256, kElemCap); }

template <class T>
void cycle(phase_hip_solver *s, Multilevel &M, int l, const Scalars *sc, const T *r, T *e);

// e = (approximate) A_l^-1 r at coarse level l
template <class T>
void coarseSolve(phase_hip_solver *s, Multilevel &M, int l, const Scalars *sc, const T *r, T *e) {
  MlLevel &L = M.lv[l];
  hipStream_t st = s->stream;
  if (!L.sol) {
    hipLaunchKernelGGL((kDenseApply<T>), dim3((L.n + 3) / 4), dim3(256), 0, st, L.n, L.inv.p, r, e, sc);
    return;
  }
  if (!s->mlKCycle || l >= s->mlKLevels) {
    cycle<T>(s, M, l, sc, r, e);
    return;
  }
  // two flexible CG iterations preconditioned by the cycle (K cycle)
  MlVecs<T> &V = vecs<T>(L);
  phase_hip_solver *c = L.sol;
  const int n = L.n, gs = spmvGrid(c), gu = gridFor(n, kRedBlock, kRedCap);
  // each update block re-reads the gs partials: keep the grid near gs
  const int gd = gridFor(n, 4 * kRedBlock, kRedCap);
  if (L.jacobi) {
    // on a Jacobi level the K cycle's updates ride on the cycle's launches:
    // first apply, A z1 with its dots, second apply (r1 formed in its first
    // launch, (z2, q1) summed in its last), the direction with A p2 and its
    // dots, the update of e
    MlLevel &N = M.lv[l + 1];
    MlVecs<T> &NV = vecs<T>(N);
    cycle<T>(s, M, l, sc, r, V.z1.p);
    hipLaunchKernelGGL((kMlSpmvDot2<T, double>), dim3(gs), dim3(kRedBlock), 0, st, n, c->ord.nSlices,
                       c->dSliceOff.p, c->dSellCol.p, c->dSell64.p, (const T *)V.z1.p, V.q1.p, r, sc, L.partials.p);
    hipLaunchKernelGGL((kMlJacFirstRestrict<T, double>), dim3((N.n + 3) / 4), dim3(256), 0, st, N.n,
                       N.sol ? N.restRow.p : nullptr, N.finerAggPtr.p, c->dSliceOff.p, c->dSellCol.p, c->dSell64.p,
                       dinvOf<T>(L), r, (const T *)V.q1.p, (const T *)V.z1.p, (const double *)L.partials.p, gs,
                       L.msc.p, V.r1.p, e, V.w.p, NV.r.p, sc);
    coarseSolve<T>(s, M, l + 1, sc, NV.r.p, NV.e.p);
    hipLaunchKernelGGL((kMlJacResidProlong<T, double, true>), dim3(gs), dim3(kRedBlock), 0, st, n, c->ord.nSlices,
                       c->dSliceOff.p, c->dSellCol.p, c->dSell64.p, N.pIdx.p, (const T *)NV.e.p, (const T *)V.w.p,
                       (const T *)V.r1.p, dinvOf<T>(L), V.z2.p, sc, (const T *)V.q1.p, L.partials2.p);
    hipLaunchKernelGGL((kMlJacDirection<T, double>), dim3(gs), dim3(kRedBlock), 0, st, n, c->ord.nSlices,
                       c->dSliceOff.p, c->dSellCol.p, c->dSell64.p, sc, (const MlScalars *)L.msc.p,
                       (const double *)L.partials2.p, gs, (const T *)V.z1.p, (const T *)V.z2.p, (const T *)V.r1.p,
                       V.t.p, L.partials.p);
    hipLaunchKernelGGL((kMlSecond<T>), dim3(gd), dim3(kRedBlock), 0, st, n, sc, L.partials.p, gs, V.t.p, e);
    return;
  }
  cycle<T>(s, M, l, sc, r, V.z1.p);
  hipLaunchKernelGGL((kMlSpmvDot2<T, double>), dim3(gs), dim3(kRedBlock), 0, st, n, c->ord.nSlices, c->dSliceOff.p,
                     c->dSellCol.p, c->dSell64.p, (const T *)V.z1.p, V.q1.p, r, sc, L.partials.p);
  hipLaunchKernelGGL((kMlFirst<T>), dim3(gd), dim3(kRedBlock), 0, st, n, sc, L.msc.p, L.partials.p, gs, V.z1.p,
                     V.q1.p, r, e, V.r1.p);
  cycle<T>(s, M, l, sc, V.r1.p, V.z2.p);
  hipLaunchKernelGGL((kMlDotZQ<T>), dim3(gu), dim3(kRedBlock), 0, st, n, sc, V.z2.p, V.q1.p, L.partials.p);
  hipLaunchKernelGGL((kMlDirection<T>), dim3(gd), dim3(kRedBlock), 0, st, n, sc, L.msc.p, L.partials.p, gu,
                     V.z1.p, V.z2.p);
  hipLaunchKernelGGL((kMlSpmvDot2<T, double>), dim3(gs), dim3(kRedBlock), 0, st, n, c->ord.nSlices, c->dSliceOff.p,
                     c->dSellCol.p, c->dSell64.p, (const T *)V.z2.p, V.q1.p, V.r1.p, sc, L.partials.p);
  hipLaunchKernelGGL((kMlSecond<T>), dim3(gd), dim3(kRedBlock), 0, st, n, sc, L.partials.p, gs, V.z2.p, e);
}

// Restriction of the residual t (finer solver order) into level l's
// right-hand side, the coarse solve, and the prolongation added to z.
template <class T>
void coarseCorrection(phase_hip_solver *s, Multilevel &M, int l, const Scalars *sc, const T *t, T *z) {
  MlLevel &L = M.lv[l];
  MlVecs<T> &V = vecs<T>(L);
  hipStream_t st = s->stream;
  hipLaunchKernelGGL((kMlRestrictRows<T>), dim3((L.n + 15) / 16), dim3(256), 0, st, L.n,
                     L.sol ? L.restRow.p : nullptr, L.finerAggPtr.p, t, V.r.p, sc);
  coarseSolve<T>(s, M, l, sc, V.r.p, V.e.p);
  hipLaunchKernelGGL((kMlProlongAdd<T>), dim3(elemGrid(L.nFiner)), dim3(256), 0, st, L.nFiner, L.pIdx.p,
                     (const T *)V.e.p, z, sc);
}

// Smoothing, coarse correction and smoothing at level `lvl` (nested solver c,
// its work vectors V): z = B r.  When `pre` is false the pre-smoothing has
// already put M^-1 r into z.
//
// On the coarse levels (c nested), where every launch costs more than its
// arithmetic, the residual and its restriction run as one launch, and so do
// the prolongation and the post-smoothing residual (into the work vector u);
// at every level the post-smoothing sweep adds its result into z as it
// finishes each row.  The arithmetic is that of the separate kernels.
template <class T>
void smoothCorrectSmooth(phase_hip_solver *s, Multilevel &M, phase_hip_solver *c, MlVecs<T> &V, int next,
                         const Scalars *sc, const T *r, T *z, bool pre, bool post = true) {
  hipStream_t st = s->stream;
  const int n = c->rank;
  const IluDev<T> L = iluDev<T>(c, facOf<T>(c));
  if (pre) launchLU<T, kProNone>(c, L, sc, r, V.y.p, (T *)nullptr, (const T *)nullptr, (const T *)nullptr, z);
  const int gs = spmvGrid(c);
  if (c != s && post) {
    MlLevel &N = M.lv[next];
    MlVecs<T> &NV = vecs<T>(N);
    hipLaunchKernelGGL((kMlResidRestrict<T, double>), dim3((N.n + 3) / 4), dim3(256), 0, st, N.n,
                       N.sol ? N.restRow.p : nullptr, N.finerAggPtr.p, c->dSliceOff.p, c->dSellCol.p,
                       c->dSell64.p, (const T *)z, r, NV.r.p, sc);
    coarseSolve<T>(s, M, next, sc, NV.r.p, NV.e.p);
    hipLaunchKernelGGL((kMlResidProlong<T, double>), dim3(gs), dim3(kRedBlock), 0, st, n, c->ord.nSlices,
                       c->dSliceOff.p, c->dSellCol.p, c->dSell64.p, N.pIdx.p, (const T *)NV.e.p, (const T *)z, r,
                       V.t.p, V.u.p, sc);
    launchLU<T, kProNone>(c, L, sc, V.t.p, V.y.p, (T *)nullptr, (const T *)nullptr, (const T *)nullptr, V.w.p,
                          nullptr, (const T *)V.u.p, z);
    return;
  }
  hipLaunchKernelGGL((kMlResid<T, double>), dim3(gs), dim3(kRedBlock), 0, st, n, c->ord.nSlices, c->dSliceOff.p,
                     c->dSellCol.p, c->dSell64.p, (const T *)z, r, V.t.p, sc);
  coarseCorrection<T>(s, M, next, sc, V.t.p, z);
  if (!post) return;
  hipLaunchKernelGGL((kMlResid<T, double>), dim3(gs), dim3(kRedBlock), 0, st, n, c->ord.nSlices, c->dSliceOff.p,
                     c->dSellCol.p, c->dSell64.p, (const T *)z, r, V.t.p, sc);
  launchLU<T, kProNone>(c, L, sc, V.t.p, V.y.p, (T *)nullptr, (const T *)nullptr, (const T *)nullptr, V.w.p,
                        nullptr, (const T *)z, z);
}

template <class T>
void cycle(phase_hip_solver *s, Multilevel &M, int l, const Scalars *sc, const T *r, T *e) {
  MlLevel &L = M.lv[l];
  if (!L.jacobi) {
    smoothCorrectSmooth<T>(s, M, L.sol, vecs<T>(L), l + 1, sc, r, e, true);
    return;
  }
  // damped Jacobi smoothing: two launches and the coarse solve
  phase_hip_solver *c = L.sol;
  MlVecs<T> &V = vecs<T>(L);
  MlLevel &N = M.lv[l + 1];
  MlVecs<T> &NV = vecs<T>(N);
  hipStream_t st = s->stream;
  hipLaunchKernelGGL((kMlJacResidRestrict<T, double>), dim3((N.n + 3) / 4), dim3(256), 0, st, N.n,
                     N.sol ? N.restRow.p : nullptr, N.finerAggPtr.p, c->dSliceOff.p, c->dSellCol.p, c->dSell64.p,
                     dinvOf<T>(L), r, V.w.p, NV.r.p, sc);
  coarseSolve<T>(s, M, l + 1, sc, NV.r.p, NV.e.p);
  hipLaunchKernelGGL((kMlJacResidProlong<T, double, false>), dim3(spmvGrid(c)), dim3(kRedBlock), 0, st, c->rank,
                     c->ord.nSlices, c->dSliceOff.p, c->dSellCol.p, c->dSell64.p, N.pIdx.p, (const T *)NV.e.p,
                     (const T *)V.w.p, r, dinvOf<T>(L), e, sc, (const T *)nullptr, (double *)nullptr);
}

}  // namespace

template <class T>
void mlFinishApply(phase_hip_solver *s, const Scalars *sc, const T *rhs, T *out) {
  Multilevel &M = *s->ml;
  if (M.lv.empty()) return;  // a system no larger than the dense limit: ILU(0) alone
  smoothCorrectSmooth<T>(s, M, s, fineVecs<T>(M), 0, sc, rhs, out, false, s->mlFinePost != 0);
}
template void mlFinishApply<double>(phase_hip_solver *, const Scalars *, const double *, double *);
template void mlFinishApply<float>(phase_hip_solver *, const Scalars *, const float *, float *);

void mlFree(phase_hip_solver *s) {
  if (!s->ml) return;
  for (MlLevel &L : s->ml->lv) freeNested(L.sol);
  delete s->ml;
  s->ml = nullptr;
}

void mlValuesChanged(phase_hip_solver *s) {
  if (s->ml) s->ml->valuesValid = s->ml->singleValid = false;
}

std::vector<int> mlLevelSizes(const phase_hip_solver *s) { return s->ml ? s->ml->sizes : std::vector<int>(); }

template <class T>
void mlResidLaunch(phase_hip_solver *s, const T *x, const T *b, T *t) {
  hipLaunchKernelGGL((kMlResid<T, double>), dim3(spmvGrid(s)), dim3(kRedBlock), 0, s->stream, s->rank,
                     s->ord.nSlices, s->dSliceOff.p, s->dSellCol.p, s->dSell64.p, x, b, t, idleScalars(s));
}
template void mlResidLaunch<double>(phase_hip_solver *, const double *, const double *, double *);
template void mlResidLaunch<float>(phase_hip_solver *, const float *, const float *, float *);

void mlCoarsest(phase_hip_solver *s, int n, double *matrix, double *inverse) {
  if (!s->ml || s->ml->lv.empty() || !s->ml->valuesValid) fail(PHASE_HIP_ESTATE, "no multigrid hierarchy");
  MlLevel &L = s->ml->lv.back();
  if (L.sol) fail(PHASE_HIP_ESTATE, "the last level is not dense");
  if (n != L.n) fail(PHASE_HIP_EINVAL, "the last level has " + std::to_string(L.n) + " rows");
  hipStream_t st = s->stream;
  // the matrix again from its CSR values (the global-memory inverse overwrites it)
  hipLaunchKernelGGL(kDenseFill, dim3(elemGrid(L.n)), dim3(256), 0, st, L.n, L.cRowPtr.p, L.cColInd.p, L.cVals.p,
                     L.dense.p);
  const size_t bytes = (size_t)L.n * L.n * sizeof(double);
  if (matrix) PH_CHECK(hipMemcpyAsync(matrix, L.dense.p, bytes, hipMemcpyDeviceToHost, st));
  if (inverse) PH_CHECK(hipMemcpyAsync(inverse, L.inv.p, bytes, hipMemcpyDeviceToHost, st));
  PH_CHECK(hipStreamSynchronize(st));
}

// First coarse level smoothed by damped Jacobi: the setting, or by default
// every level under CG and none under BiCGStab (on the non-symmetric
// ghost-cell pressure rows at 4M cells, BiCGStab with Jacobi-smoothed coarse
// levels did not reach 1e-8 in 8,000 iterations; with ILU(0) ones it took 5,806)
int jacobiFromOf(const phase_hip_solver *s) { return s->mlJacobiFrom >= 0 ? s->mlJacobiFrom : s->cg ? 0 : 1 << 20; }

void mlEnsureHierarchy(phase_hip_solver *s) {
  if (s->ml && s->ml->jacobiFrom != jacobiFromOf(s)) mlFree(s);
  if (s->ml) return;
  if (!s->havePattern) fail(PHASE_HIP_ESTATE, "no pattern");
  auto *M = new Multilevel();
  M->jacobiFrom = jacobiFromOf(s);
  try {
    const Ordering *fine = &s->ord;
    std::vector<double> xy = s->cellXY;
    int nFiner = s->rank;
    M->sizes.push_back(nFiner);
    hipStream_t st = s->stream;
    for (int level = 0; nFiner > s->mlDenseRows; ++level) {
      CoarseLevel c;
      buildCoarse(*fine, xy.empty() ? nullptr : xy.data(), c);
      // the restriction kernel takes one aggregate per wavefront
      for (size_t a = 0; a + 1 < fine->aggPtr.size(); ++a)
        if (fine->aggPtr[a + 1] - fine->aggPtr[a] > kLanes)
          fail(PHASE_HIP_EINVAL, "multigrid: aggregate of more than 64 rows");
      if (c.n >= nFiner) fail(PHASE_HIP_EINVAL, "multigrid: aggregation did not coarsen level " + std::to_string(level));
      M->lv.emplace_back();
      MlLevel &L = M->lv.back();
      L.n = c.n;
      L.nFiner = nFiner;
      L.nSlots = (int)c.colInd.size();
      L.galPtr.upload(c.galPtr, st);
      L.galSrc.upload(c.galSrc.empty() ? std::vector<int>{0} : c.galSrc, st);
      L.finerAggPtr.upload(fine->aggPtr, st);
      std::vector<int> pIdx(nFiner);
      bool dense = c.n <= s->mlDenseRows;
      if (!dense) {
        L.sol = makeNested(s);
        L.sol->blockJacobi = level >= s->mlBlockJacobiFrom;
        L.jacobi = level >= M->jacobiFrom;
        L.sol->iluLimits = L.jacobi ? 0 : 1;
        // coarse rows couple to more rows than fine ones: when a row of an
        // aggregate would couple to more rows of its own block than the sweep
        // kernels take, smaller aggregates are taken at this level; a level
        // no ordering fits is inverted densely when it is small enough
        for (int br = s->mlCoarseRows;; br /= 2) {
          L.sol->blockRows = br;
          try {
            solverSetPattern(L.sol, c.n, c.rowPtr.data(), c.colInd.data(), c.xy.empty() ? nullptr : c.xy.data());
            break;
          } catch (const HipError &e) {
            if (br > 2 && e.code == PHASE_HIP_EINVAL) continue;
            if (e.code != PHASE_HIP_EINVAL || c.n > kDenseMax) throw;
            freeNested(L.sol);
            L.sol = nullptr;
            dense = true;
            break;
          }
        }
      }
      if (dense) {
        if (c.n > kDenseMax) fail(PHASE_HIP_EINVAL, "multigrid: dense level too large");
        L.cRowPtr.upload(c.rowPtr, st);
        L.cColInd.upload(c.colInd, st);
        L.cVals.alloc(std::max(1, L.nSlots));
        L.dense.alloc((size_t)c.n * c.n);
        L.inv.alloc((size_t)c.n * c.n);
        for (int i = 0; i < nFiner; ++i) pIdx[i] = c.aggOfRow[i];
      } else {
        L.restRow.upload(L.sol->ord.perm, st);
        for (int i = 0; i < nFiner; ++i) pIdx[i] = L.sol->ord.iperm[c.aggOfRow[i]];
        L.msc.alloc(1);
        L.partials.alloc(2 * (size_t)std::max(spmvGrid(L.sol), kRedCap));
        L.partials2.alloc((size_t)std::max(spmvGrid(L.sol), kRedCap));
      }
      L.pIdx.upload(pIdx, st);
      M->sizes.push_back(c.n);
      if (!L.sol) break;
      fine = &L.sol->ord;
      xy = c.xy;
      nFiner = c.n;
    }
    PH_CHECK(hipStreamSynchronize(st));
  } catch (...) {
    for (MlLevel &L : M->lv) freeNested(L.sol);
    delete M;
    throw;
  }
  s->ml = M;
}

void mlEnsureValues(phase_hip_solver *s, bool single) {
  Multilevel &M = *s->ml;
  hipStream_t st = s->stream;
  if (!M.valuesValid) {
    const double *finer = s->dVals.p;
    for (MlLevel &L : M.lv) {
      if (L.sol) {
        hipLaunchKernelGGL(kGalerkin, dim3(elemGrid(L.nSlots)), dim3(256), 0, st, L.nSlots, L.galPtr.p, L.galSrc.p,
                           finer, L.sol->dVals.p);
        solverValuesChanged(L.sol);
        if (L.jacobi) {
          L.dinv64.alloc(L.n);
          hipLaunchKernelGGL(kMlDiagInv, dim3(elemGrid(L.n)), dim3(256), 0, st, L.n, L.sol->dSliceOff.p,
                             L.sol->dSellCol.p, L.sol->dSell64.p, s->mlJacobiWeight, L.dinv64.p);
        } else {
          ensureFactor64(L.sol);
        }
        finer = L.sol->dVals.p;
      } else {
        hipLaunchKernelGGL(kGalerkin, dim3(elemGrid(L.nSlots)), dim3(256), 0, st, L.nSlots, L.galPtr.p, L.galSrc.p,
                           finer, L.cVals.p);
        hipLaunchKernelGGL(kDenseFill, dim3(elemGrid(L.n)), dim3(256), 0, st, L.n, L.cRowPtr.p, L.cColInd.p,
                           L.cVals.p, L.dense.p);
        denseInverse(st, L.n, L.dense.p, L.inv.p, s->dErr.p, s->mlDensePivot != 0);
        int e = 0;
        PH_CHECK(hipMemcpyAsync(&e, s->dErr.p, sizeof(int), hipMemcpyDeviceToHost, st));
        PH_CHECK(hipStreamSynchronize(st));
        if (e & 4) {
          PH_CHECK(hipMemsetAsync(s->dErr.p, 0, sizeof(int), st));
          fail(PHASE_HIP_EBREAKDOWN, "multigrid: singular coarsest matrix");
        }
      }
    }
    M.valuesValid = true;
    M.singleValid = false;
  }
  // work vectors: FP64 always (a mixed solve may finish in FP64), FP32 for mixed solves
  if (M.lv.empty()) return;
  const int n = s->rank;
  M.f64.allocCycle(n);
  for (MlLevel &L : M.lv) L.v64.alloc(L.n);
  if (single) {
    M.f32.allocCycle(n);
    for (MlLevel &L : M.lv) L.v32.alloc(L.n);
    if (!M.singleValid) {
      for (MlLevel &L : M.lv)
        if (L.jacobi) {
          L.dinv32.alloc(L.n);
          hipLaunchKernelGGL((kConvert<double, float>), dim3(elemGrid(L.n)), dim3(256), 0, st, L.n, L.dinv64.p,
                             L.dinv32.p);
        } else if (L.sol) {
          ensureSingle(L.sol);
        }
      M.singleValid = true;
    }
  }
  PH_CHECK(hipGetLastError());
}

}  // namespace phase
