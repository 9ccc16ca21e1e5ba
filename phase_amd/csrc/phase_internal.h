// Internal declarations of libphase_hip: device buffers, the solver ordering
// and the kernel launchers shared by solver.hip and mesh.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/phase_hip.h"

namespace phase {

struct HipError {
  int code;
  std::string msg;
};

#define PH_CHECK(expr)                                                                   \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      throw ::phase::HipError{PHASE_HIP_EDEVICE,                                         \
                              std::string(#expr) + ": " + hipGetErrorString(_e)};        \
  } while (0)

[[noreturn]] inline void fail(int code, const std::string &msg) { throw HipError{code, msg}; }

// Owning device allocation (hipMalloc / hipFree), movable.
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count) PH_CHECK(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  void upload(const std::vector<T> &h, hipStream_t s) {
    alloc(h.size());
    if (n) PH_CHECK(hipMemcpyAsync(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void zero(hipStream_t s) {
    if (n) PH_CHECK(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
};

constexpr int kLanes = 64;  // rows per ILU block and per SpMV slice (one wavefront)

// Block metadata of the ILU ordering (one wavefront per block).
struct BlockInfo {
  int rowStart;   // first solver row
  int nRows;      // <= 64
  int lOff, kL;   // L slots: lOff + e*64 + lane, e < kL
  int uOff, kU;   // U slots
  int nLev, nULev;  // in-block levels of the forward / backward sweeps
  int predBeg, predEnd;  // blocks the forward sweep waits for (pred[])
  int succBeg, succEnd;  // blocks the backward sweep waits for (succ[])
};

// Per-row ILU metadata: counts of L / U entries and the in-/off-block split.
struct RowInfo {
  uint8_t nL, nLoff, nU, nUin;
  uint8_t lev, ulev, pad0, pad1;
};

// Per-lane sweep record, indexed by block * 64 + lane so that a sweep reads it
// in parallel with the block record.  row = solver row (-1: idle lane); info
// packs the counts and in-block levels of the row:
//   bits 0-3 in-block L, 4-9 off-block L, 10-13 in-block U, 14-17 off-block U,
//   18-23 forward level, 24-29 backward level.
struct LaneRec {
  int row;
  unsigned info;
};
inline __host__ __device__ unsigned packLane(int nLin, int nLoff, int nUin, int nUoff, int lev, int ulev) {
  return (unsigned)nLin | ((unsigned)nLoff << 4) | ((unsigned)nUin << 10) | ((unsigned)nUoff << 14) |
         ((unsigned)lev << 18) | ((unsigned)ulev << 24);
}

// Host-built ordering of one sparsity pattern.  "compact" = the caller's CSR
// with empty slots dropped, original numbering and in-row order.
struct Ordering {
  int n = 0;
  int nnz = 0;                       // compact entries
  std::vector<int> cRowPtr, cCol;    // compact pattern, original numbering
  std::vector<int> compactSrc;       // compact entry -> caller slot (padded index)
  std::vector<int> perm, iperm;      // solver row k <-> original row perm[k]
  int nBlocks = 0, nColours = 0;
  int regWidth = 8;                  // sweep kernels' per-row register width (4 when every count fits)
  std::vector<int> colourPtr;         // blocks of colour c: [colourPtr[c], colourPtr[c+1])
  // aggregates (the rows a block packs, one aggregate at a time): aggregate a
  // holds solver rows [aggPtr[a], aggPtr[a+1]), numbered in solver order
  std::vector<int> aggPtr;
  std::vector<BlockInfo> blocks;
  std::vector<int> pred, succ;
  std::vector<RowInfo> rows;          // per solver row
  std::vector<LaneRec> lanes;         // per block * 64 + lane
  // SELL-64 SpMV layout over solver rows: slice s covers rows [64s, 64s+64)
  int nSlices = 0;
  std::vector<int> sliceOff;          // nSlices+1, in units of slots
  std::vector<int> sellCol;           // solver column or -1
  std::vector<int> sellSrc;           // compact entry index or -1
  // ILU factor layout: fac[] = [L slots | U slots | diagonal (n)]
  int nLSlots = 0, nUSlots = 0;
  std::vector<int> lCol, uCol;        // solver column, -1 padding
  std::vector<int> facSrc;            // per fac entry: compact entry index or -1
  std::vector<int> updPtr;            // per L slot (nLSlots+1)
  std::vector<int> updSrc, updDst;    // fac indices
};

// cellXY (optional): centroid (x, y) of each row's cell, interleaved; with it
// a scalar equation on a 2-D mesh gets three-colour brick aggregates.
// blockJacobi: one colour, and the ILU(0) factor of the block-diagonal part
// only (couplings between blocks dropped): one sweep launch per apply.
// iluLimits false: the ordering's ILU(0) will not be applied (a multigrid
// level smoothed by Jacobi), so rows wider than the sweep kernels' lane
// records are accepted.
void buildOrdering(Ordering &o, int n, const int *rowPtr, const int *colInd, int blockRows,
                   const double *cellXY = nullptr, bool blockJacobi = false, bool iluLimits = true);

// One coarser level of the aggregation multigrid hierarchy, built from the
// aggregates of a finer level's ordering (unsmoothed aggregation: the
// prolongation P has one 1 per row, in the column of the row's aggregate, and
// the coarse matrix is the Galerkin product P^T A P).  Coarse row a is
// aggregate a of the finer ordering.
struct CoarseLevel {
  int n = 0;
  std::vector<int> rowPtr, colInd;   // coarse pattern (compact CSR, ascending columns)
  std::vector<double> xy;            // aggregate centroids (interleaved), empty without finer centroids
  std::vector<int> galPtr, galSrc;   // coarse slot q = sum of finer caller slots galSrc[galPtr[q] .. galPtr[q+1])
  std::vector<int> aggOfRow;         // finer solver row -> coarse row
};
// fineXY: the finer level's centroids in its caller numbering, or null.
void buildCoarse(const Ordering &fine, const double *fineXY, CoarseLevel &c);

}  // namespace phase

namespace phase {

// Device-resident scalars of one BiCGStab run.  Written only by the last
// workgroup of each reduction kernel (deterministic fixed-order sums).
struct Scalars {
  double rho, alpha, omega, beta;
  double rr, tol2;        // ||r||^2 of the recursive residual and the stop threshold
  double rv, ts, tt;      // last (r0,v), (t,s), (t,t)
  double rrPrev;          // ||r||^2 one iteration earlier (fallback monitor)
  int done;               // 0 running, 1 converged, 2 breakdown, 3 max iterations, 4 stagnation
  int iters, maxIters, pad;
  double minRatio;        // learned mode: smallest accepted ||r_k||/||r_{k-1}|| ... (<=0 disables)
  double rhoNext;         // CG: (r, z) of the newest z, taken over as rho by the next alpha
};

enum { kDoneRunning = 0, kDoneConverged = 1, kDoneBreakdown = 2, kDoneMaxIters = 3,
       kDoneStagnation = 4 };

}  // namespace phase
