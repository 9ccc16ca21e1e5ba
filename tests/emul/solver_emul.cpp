// CPU emulation of the device solver's layout and kernels (test-only).
//
// Builds the product's ordering (phase_amd/csrc/ordering.cpp) and replays the
// SELL-64 SpMV, the ILU(0) factorisation and both triangular sweeps exactly as
// the gfx950 kernels in solver_kernels.h do -- block by block in ticket order,
// level by level inside a block -- while checking every array index against
// its bounds and every dependency against the schedule (a predecessor block
// must already be finished, an in-block dependency must sit at a lower level).
// The results are compared with the oracle by tests/test_ordering_emul.py.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../phase_amd/csrc/phase_internal.h"

using namespace phase;

namespace {
struct Bad {
  std::string msg;
};
#define REQ(c, m) \
  do {            \
    if (!(c)) throw Bad{m}; \
  } while (0)
}  // namespace

namespace {
bool gBlockJacobi = false;
}
// The next emul_run builds block-Jacobi orderings (one colour, factor of the
// block-diagonal part).
extern "C" void emul_set_block_jacobi(int on) { gBlockJacobi = on != 0; }

extern "C" int emul_run(int n, const int *rowPtr, const int *colInd, const double *vals, int blockRows,
                        int *perm, int *facRowPtr, int *facCol, double *facVal, int cap, const double *rhsIn,
                        double *y, double *x, const double *xin, double *spmvOut, int *info, char *err,
                        int errLen, const double *cellXY) {
  try {
    Ordering o;
    try {
      buildOrdering(o, n, rowPtr, colInd, blockRows, cellXY, gBlockJacobi);
    } catch (const HipError &e) {
      throw Bad{e.msg};
    }
    const int nslots = rowPtr[n];
    const int nb = o.nBlocks;
    info[0] = nb;
    info[1] = o.nColours;
    info[2] = o.nLSlots;
    info[3] = o.nUSlots;
    int maxLev = 0, maxULev = 0;
    for (auto &B : o.blocks) {
      maxLev = std::max(maxLev, B.nLev);
      maxULev = std::max(maxULev, B.nULev);
    }
    info[4] = maxLev;
    info[5] = maxULev;
    int maxIn = 0;  // widest in-block L / in-block U / off-block U count of a row
    for (const RowInfo &r : o.rows) maxIn = std::max({maxIn, r.nL - r.nLoff, (int)r.nUin, r.nU - r.nUin});
    info[6] = maxIn;
    for (int k = 0; k < n; ++k) perm[k] = o.perm[k];
    // --- value gathers (kGather)
    const size_t nSell = o.sellCol.size();
    REQ(o.sellSrc.size() == nSell, "sellSrc size");
    std::vector<double> sell(nSell);
    for (size_t i = 0; i < nSell; ++i) {
      REQ(o.sellSrc[i] < nslots, "sellSrc out of range");
      sell[i] = o.sellSrc[i] >= 0 ? vals[o.sellSrc[i]] : 0.;
    }
    const size_t nFac = (size_t)o.nLSlots + o.nUSlots + n;
    REQ(o.facSrc.size() == nFac, "facSrc size");
    std::vector<double> fac(nFac);
    for (size_t i = 0; i < nFac; ++i) {
      REQ(o.facSrc[i] < nslots, "facSrc out of range");
      fac[i] = o.facSrc[i] >= 0 ? vals[o.facSrc[i]] : 0.;
    }
    // --- SpMV (kSpmv)
    REQ((int)o.sliceOff.size() == o.nSlices + 1, "sliceOff size");
    for (int k = 0; k < n; ++k) {
      int s = k / 64, lane = k % 64;
      int off = o.sliceOff[s], K = (o.sliceOff[s + 1] - off) / 64;
      REQ((o.sliceOff[s + 1] - off) % 64 == 0, "slice not a multiple of 64");
      double acc = 0.;
      for (int e = 0; e < K; ++e) {
        size_t idx = off + e * 64 + lane;
        REQ(idx < nSell, "sell index out of range");
        int c = o.sellCol[idx];
        REQ(c < n, "sell column out of range");
        if (c >= 0) acc = std::fma(sell[idx], xin[c], acc);
      }
      spmvOut[k] = acc;
    }
    const int dbase = o.nLSlots + o.nUSlots;
    auto lslot = [&](const BlockInfo &B, int e, int lane) {
      int s = B.lOff + e * 64 + lane;
      REQ(s >= 0 && s < o.nLSlots, "L slot out of range");
      return s;
    };
    auto uslot = [&](const BlockInfo &B, int e, int lane) {
      int s = B.uOff + e * 64 + lane;
      REQ(s >= 0 && s < o.nUSlots, "U slot out of range");
      return s;
    };
    // per-colour launches: every predecessor block of a block lies in an
    // earlier colour, every successor in a later one
    REQ((int)o.colourPtr.size() == o.nColours + 1 && o.colourPtr[0] == 0 && o.colourPtr[o.nColours] == nb,
        "colourPtr");
    std::vector<int> colourOf(nb, -1);
    for (int c = 0; c < o.nColours; ++c)
      for (int b = o.colourPtr[c]; b < o.colourPtr[c + 1]; ++b) colourOf[b] = c;
    for (int b = 0; b < nb; ++b) {
      const BlockInfo &B = o.blocks[b];
      for (int q = B.predBeg; q < B.predEnd; ++q) REQ(colourOf[o.pred[q]] < colourOf[b], "pred not in an earlier colour");
      for (int q = B.succBeg; q < B.succEnd; ++q) REQ(colourOf[o.succ[q]] > colourOf[b], "succ not in a later colour");
    }
    // per-lane sweep records agree with the row records
    REQ(o.lanes.size() == (size_t)nb * 64, "lane record count");
    for (int b = 0; b < nb; ++b) {
      const BlockInfo &B = o.blocks[b];
      for (int lane = 0; lane < 64; ++lane) {
        const LaneRec &L = o.lanes[(size_t)b * 64 + lane];
        if (lane >= B.nRows) {
          REQ(L.row == -1, "idle lane with a row");
          continue;
        }
        const int k = B.rowStart + lane;
        const RowInfo &ri = o.rows[k];
        REQ(L.row == k, "lane row");
        REQ((int)(L.info & 15u) == ri.nL - ri.nLoff && (int)((L.info >> 4) & 63u) == ri.nLoff &&
                (int)((L.info >> 10) & 15u) == ri.nUin && (int)((L.info >> 14) & 15u) == ri.nU - ri.nUin &&
                (int)((L.info >> 18) & 63u) == ri.lev && (int)((L.info >> 24) & 63u) == ri.ulev,
            "lane info packing");
      }
    }
    std::vector<char> doneBlk(nb, 0), rowDone(n, 0);
    // --- factorisation (kIluFactor), blocks in ticket order
    for (int b = 0; b < nb; ++b) {
      const BlockInfo &B = o.blocks[b];
      REQ(B.nRows >= 1 && B.nRows <= 64, "block size");
      REQ(B.predEnd - B.predBeg <= 64, "too many preds");
      for (int q = B.predBeg; q < B.predEnd; ++q) {
        REQ(o.pred[q] < b, "pred block not earlier");
        REQ(doneBlk[o.pred[q]], "pred block not finished");
      }
      for (int lev = 0; lev < B.nLev; ++lev) {
        for (int lane = 0; lane < B.nRows; ++lane) {
          const int row = B.rowStart + lane;
          const RowInfo &ri = o.rows[row];
          if (ri.lev != lev) continue;
          for (int e = 0; e < ri.nL; ++e) {
            const int s = lslot(B, e, lane);
            const int k = o.lCol[s];
            REQ(k >= 0 && k < row, "L column not below diagonal");
            if (e < ri.nLoff) REQ(k < B.rowStart, "off-block L entry inside block");
            else REQ(k >= B.rowStart && o.rows[k].lev < lev, "in-block L dependency not at a lower level");
            REQ(rowDone[k], "L dependency row not finished");
            const double lik = fac[s] / fac[dbase + k];
            fac[s] = lik;
            REQ(o.updPtr[s] <= o.updPtr[s + 1], "updPtr");
            for (int u = o.updPtr[s]; u < o.updPtr[s + 1]; ++u) {
              const int dst = o.updDst[u], src = o.updSrc[u];
              REQ(dst >= 0 && (size_t)dst < nFac && src >= o.nLSlots && src < dbase, "update index");
              fac[dst] = std::fma(-lik, fac[src], fac[dst]);
            }
          }
          rowDone[row] = 1;
        }
      }
      for (int lane = 0; lane < B.nRows; ++lane) REQ(rowDone[B.rowStart + lane], "row left unprocessed");
      doneBlk[b] = 1;
    }
    // export factor (solver order, sorted columns)
    int w = 0;
    facRowPtr[0] = 0;
    for (int b = 0; b < nb; ++b) {
      const BlockInfo &B = o.blocks[b];
      for (int lane = 0; lane < B.nRows; ++lane) {
        const int k = B.rowStart + lane;
        const RowInfo &ri = o.rows[k];
        REQ(w + ri.nL + ri.nU + 1 <= cap, "factor capacity");
        for (int e = 0; e < ri.nL; ++e) { int s = lslot(B, e, lane); facCol[w] = o.lCol[s]; facVal[w++] = fac[s]; }
        facCol[w] = k;
        facVal[w++] = fac[dbase + k];
        for (int e = 0; e < ri.nU; ++e) {
          int s = uslot(B, e, lane);
          facCol[w] = o.uCol[s];
          facVal[w++] = fac[o.nLSlots + s];
        }
        facRowPtr[k + 1] = w;
      }
    }
    // --- forward sweep (kLSolve)
    std::fill(doneBlk.begin(), doneBlk.end(), 0);
    std::fill(rowDone.begin(), rowDone.end(), 0);
    for (int b = 0; b < nb; ++b) {
      const BlockInfo &B = o.blocks[b];
      for (int q = B.predBeg; q < B.predEnd; ++q) REQ(doneBlk[o.pred[q]], "fwd pred not finished");
      std::vector<double> acc(64, 0.), lds(64, 0.);
      for (int lane = 0; lane < B.nRows; ++lane) {
        const int row = B.rowStart + lane;
        const RowInfo &ri = o.rows[row];
        acc[lane] = rhsIn[row];
        for (int e = 0; e < ri.nLoff; ++e) {
          const int s = lslot(B, e, lane);
          REQ(rowDone[o.lCol[s]], "fwd off-block value not ready");
          acc[lane] = std::fma(-fac[s], y[o.lCol[s]], acc[lane]);
        }
      }
      for (int lev = 0; lev < B.nLev; ++lev) {
        for (int lane = 0; lane < B.nRows; ++lane) {
          const int row = B.rowStart + lane;
          const RowInfo &ri = o.rows[row];
          if (ri.lev != lev) continue;
          for (int e = ri.nLoff; e < ri.nL; ++e) {
            const int s = lslot(B, e, lane);
            const int li = o.lCol[s] - B.rowStart;
            REQ(li >= 0 && li < lane && rowDone[o.lCol[s]], "fwd in-block value not ready");
            acc[lane] = std::fma(-fac[s], lds[li], acc[lane]);
          }
          lds[lane] = acc[lane];
          y[row] = acc[lane];
        }
        for (int lane = 0; lane < B.nRows; ++lane)
          if (o.rows[B.rowStart + lane].lev == lev) rowDone[B.rowStart + lane] = 1;
      }
      doneBlk[b] = 1;
    }
    // --- backward sweep (kUSolve), blocks in reverse
    std::fill(doneBlk.begin(), doneBlk.end(), 0);
    std::fill(rowDone.begin(), rowDone.end(), 0);
    for (int t = 0; t < nb; ++t) {
      const int b = nb - 1 - t;
      const BlockInfo &B = o.blocks[b];
      for (int q = B.succBeg; q < B.succEnd; ++q) {
        REQ(o.succ[q] > b, "succ block not later");
        REQ(doneBlk[o.succ[q]], "bwd succ not finished");
      }
      std::vector<double> acc(64, 0.), lds(64, 0.), soff(64, 0.);
      for (int lane = 0; lane < B.nRows; ++lane) {
        const int row = B.rowStart + lane;
        const RowInfo &ri = o.rows[row];
        acc[lane] = y[row];
        // later-block terms: one sum, formed before the levels (kUSolve)
        REQ(ri.nU - ri.nUin <= 15, "too many off-block U entries");
        for (int e = ri.nUin; e < ri.nU; ++e) {
          const int s = uslot(B, e, lane);
          REQ(o.uCol[s] >= B.rowStart + B.nRows && rowDone[o.uCol[s]], "bwd off-block value not ready");
          soff[lane] = std::fma(fac[o.nLSlots + s], x[o.uCol[s]], soff[lane]);
        }
      }
      for (int lev = 0; lev < B.nULev; ++lev) {
        for (int lane = 0; lane < B.nRows; ++lane) {
          const int row = B.rowStart + lane;
          const RowInfo &ri = o.rows[row];
          if (ri.ulev != lev) continue;
          for (int e = 0; e < ri.nUin; ++e) {
            const int s = uslot(B, e, lane);
            const int li = o.uCol[s] - B.rowStart;
            REQ(li > lane && li < B.nRows && rowDone[o.uCol[s]], "bwd in-block value not ready");
            acc[lane] = std::fma(-fac[o.nLSlots + s], lds[li], acc[lane]);
          }
          acc[lane] = (acc[lane] - soff[lane]) * (1.0 / fac[dbase + row]);
          lds[lane] = acc[lane];
          x[row] = acc[lane];
        }
        for (int lane = 0; lane < B.nRows; ++lane)
          if (o.rows[B.rowStart + lane].ulev == lev) rowDone[B.rowStart + lane] = 1;
      }
      doneBlk[b] = 1;
    }
    return 0;
  } catch (const Bad &e) {
    std::snprintf(err, errLen, "%s", e.msg.c_str());
    return 1;
  }
}

// Aggregation hierarchy of the multigrid preconditioner, built with the
// product's host code (buildOrdering per level, buildCoarse between levels)
// until a level has at most denseMax rows.  Writes, per level l < last: n,
// perm[n], the aggregate count na, aggPtr[na + 1], then the coarse pattern
// rowPtr[na + 1], nnz, colInd[nnz], galPtr[nnz + 1], galSrc[galPtr[nnz]];
// the file ends with the last level's n and a 0.  All int32.
extern "C" int emul_hierarchy(int n, const int *rowPtr, const int *colInd, const double *cellXY, int fineRows,
                              int coarseRows, int denseMax, const char *outPath, char *err, int errLen) {
  try {
    FILE *f = std::fopen(outPath, "wb");
    REQ(f != nullptr, "cannot open output");
    auto put = [&](const int *p, size_t k) { std::fwrite(p, sizeof(int), k, f); };
    std::vector<int> rp(rowPtr, rowPtr + n + 1), ci(colInd, colInd + rowPtr[n]);
    std::vector<double> xy;
    if (cellXY) xy.assign(cellXY, cellXY + 2 * (size_t)n);
    int level = 0;
    while (n > denseMax) {
      Ordering o;
      CoarseLevel c;
      try {
        buildOrdering(o, n, rp.data(), ci.data(), level == 0 ? fineRows : coarseRows, xy.empty() ? nullptr : xy.data());
        buildCoarse(o, xy.empty() ? nullptr : xy.data(), c);
      } catch (const HipError &e) {
        std::fclose(f);
        throw Bad{"level " + std::to_string(level) + ": " + e.msg};
      }
      REQ(c.n < n, "aggregation did not coarsen");
      // every finer compact entry lands in exactly one coarse slot
      std::vector<int> seen(rp[n], 0);
      for (int s : c.galSrc) {
        REQ(s >= 0 && s < rp[n] && ci[s] >= 0, "Galerkin source out of range");
        ++seen[s];
      }
      for (int s = 0; s < rp[n]; ++s) REQ(seen[s] == (ci[s] >= 0 ? 1 : 0), "Galerkin map does not cover the entries once");
      put(&n, 1);
      put(o.perm.data(), n);
      const int na = c.n;
      put(&na, 1);
      put(o.aggPtr.data(), na + 1);
      put(c.rowPtr.data(), na + 1);
      const int nnz = (int)c.colInd.size();
      put(&nnz, 1);
      put(c.colInd.data(), nnz);
      put(c.galPtr.data(), nnz + 1);
      put(c.galSrc.data(), c.galSrc.size());
      n = na;
      rp = c.rowPtr;
      ci = c.colInd;
      xy = c.xy;
      ++level;
    }
    put(&n, 1);
    const int zero = 0;
    put(&zero, 1);
    std::fclose(f);
    return level;
  } catch (const Bad &b) {
    std::snprintf(err, errLen, "%s", b.msg.c_str());
    return -1;
  }
}
