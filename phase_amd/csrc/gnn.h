// Learned correction network (message passing on the matrix graph) -- device
// side.  Its mathematics (features, layers, loss) is written out in the
// header of gnn.hip; training and inference kernels use BF16 MFMA
// (16x16x32, FP32 accumulate) for every dense layer.
#pragma once

#include <cstdint>
#include <vector>

#include "phase_internal.h"

struct phase_hip_solver;

namespace phase {

constexpr int kGnnH = 64;      // hidden width
constexpr int kGnnDin = 4;     // input features
constexpr int kGnnMaxL = 8;    // maximum message-passing layers

// Flat FP32 parameter layout (also the checkpoint format):
//   W_in [4][64], b_in [64],
//   per layer l: Wsm [128][64] (rows 0..63 = W_s, 64..127 = W_m), Wg [64][64], b [64],
//   w_out [64], b_out [1]
struct GnnLayout {
  int L = 6;
  size_t inW() const { return 0; }
  size_t inB() const { return kGnnDin * kGnnH; }
  size_t layer(int l) const { return inB() + kGnnH + (size_t)l * layerSize(); }
  static size_t layerSize() { return 128 * kGnnH + kGnnH * kGnnH + kGnnH; }
  size_t wsm(int l) const { return layer(l); }
  size_t wg(int l) const { return layer(l) + 128 * kGnnH; }
  size_t b(int l) const { return layer(l) + 128 * kGnnH + kGnnH * kGnnH; }
  size_t outW() const { return layer(L); }
  size_t outB() const { return outW() + kGnnH; }
  size_t total() const { return outB() + 1; }
};

struct GnnState {
  GnnLayout lay;
  int n = 0;
  bool initialised = false, trained = false, symmetric = false;
  int step = 0;
  DBuf<float> params, grad, adamM, adamV;
  DBuf<__bf16> wT;     // per layer [64 out][128 k] (forward)
  DBuf<__bf16> wN;     // per layer [128 k][64 out] (backward data)
  DBuf<float> sd, rowsum, X, out, dOut;
  DBuf<float> resC, resQ, resDq, resG;  // residual objective: out / sd, A c, dloss/dq, A dloss/dq
  DBuf<__bf16> h;      // (L+1) x n x 64
  DBuf<__bf16> agg, pre;  // L x n x 64
  DBuf<__bf16> pre0, dpre;
  DBuf<float> dh, dSA;    // n x 64, n x 128
  DBuf<float> gmean;      // L x 64
  DBuf<float> cvec;       // L x 64 (b + g Wg)
  DBuf<float> partials;
  DBuf<double> scal;      // [0] rms scale s, [1] loss
  DBuf<unsigned> counter;
  // captured (residual, correction) pairs, solver order, FP32
  int captureEvery = 0, maxPairs = 0, nPairs = 0;
  std::vector<int> pendingSnap;           // pair slots of the running solve
  DBuf<float> pairR, pairE;               // maxPairs x n
  DBuf<double> snapX;                     // x snapshots of the running solve
  std::vector<float> hostParamsCache;
};

void gnnFree(phase_hip_solver *s);
// Network ready for the current matrix (FP32 copy, graph scaling, BF16 weights).
void gnnPrepare(phase_hip_solver *s);
// Called by the FP64 solver loop: snapshot (r, x) when capture is on; the
// captured corrections e = x_final - x are kept only if the solve converged.
void gnnCaptureIter(phase_hip_solver *s, const double *r, const double *x, int iter);
void gnnCaptureFinish(phase_hip_solver *s, const double *xFinal, bool converged);
// e (solver order, FP64) = network correction for residual r64 (solver order).
void gnnApply(phase_hip_solver *s, const double *r64, double *e64);

}  // namespace phase
