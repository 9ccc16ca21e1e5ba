// fv:: operators and CrsEquation algebra on the device.
//
// A phase_hip_equation is FiniteVolumeEquation<T> / CrsEquation
// (src/Math/CrsEquation.cpp, src/2D/Unstructured/FiniteVolume/Equation/
// FiniteVolumeEquation.tpp) held in HBM: rank = ncomp * nCells rows in the
// IndexMap block layout (row = comp * nCells + cell, IndexMap.cpp:31-36), each
// row a list of up to `stride` (column, value) pairs in first-touch order plus
// the stored rhs (A x + rhs = 0).
//
// Operators run one thread per cell and replay the reference's per-cell face
// loop, so every row has the reference's column order and every sum its
// summation order (bit-identical with -ffp-contract=off).  A row is built in
// registers with CrsEquation::addCoeff's first-touch rule (CrsEquation.cpp:
// 113-131).  operator+= / -= (CrsEquation.cpp:185-275) rebuilds each row the
// same way: the left rows' non-zero entries, then the right rows' non-zero
// entries accumulated into a matching column or appended.
//
// Field kernels: src::div / src::src (Source.cpp:5-25, 77-95), the face and
// cell gradients (ScalarGradient.cpp:34-74) and face interpolation with the
// boundary rules (FiniteVolumeField.tpp:129-182, VectorFiniteVolumeField.cpp:
// 140-160).  Vector fields are interleaved (x0, y0, x1, y1, ...).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "equation.h"
#include "mesh.h"
#include "solver.h"
#include "solver_kernels.h"

using namespace phase;

namespace {

constexpr int kRowCap = 16;  // register row capacity: diagonal + 14 neighbours + one cross-component entry

// One row under construction, CrsEquation::addCoeff semantics: the first touch
// of a column claims the next slot, later touches accumulate in call order.
struct RowB {
  int col[kRowCap];
  double val[kRowCap];
  int n;
  bool over;
  __device__ void init() {
    n = 0;
    over = false;
#pragma unroll
    for (int j = 0; j < kRowCap; ++j) {
      col[j] = -1;
      val[j] = 0.;
    }
  }
  __device__ void add(int c, double v) {
    bool done = false;
#pragma unroll
    for (int j = 0; j < kRowCap; ++j)
      if (!done && j < n && col[j] == c) {
        val[j] += v;
        done = true;
      }
    if (done) return;
    if (n >= kRowCap) {
      over = true;
      return;
    }
#pragma unroll
    for (int j = 0; j < kRowCap; ++j)
      if (j == n) {
        col[j] = c;
        val[j] = v;
      }
    ++n;
  }
  __device__ void store(int row, int stride, int *cols, double *vals, int *cnt, int *err) const {
    if (over || n > stride) {
      atomicOr(err, 4);
      return;
    }
#pragma unroll
    for (int j = 0; j < kRowCap; ++j)
      if (j < stride) {
        cols[(size_t)row * stride + j] = j < n ? col[j] : -1;
        vals[(size_t)row * stride + j] = j < n ? val[j] : 0.;
      }
    cnt[row] = n;
  }
};

struct MeshDev {
  int nCells;
  const double *vol, *cc, *fc, *sf;
  const int *iptr, *iface, *inb, *bptr, *bface;
  const double *idot, *ims, *bdot, *bms;
};

MeshDev meshDev(const phase_hip_mesh *m) {
  return MeshDev{m->nCells, m->dVol.p, m->dCC.p, m->dFC.p, m->dSf.p, m->dIptr.p, m->dIface.p, m->dInb.p,
                 m->dBptr.p, m->dBface.p, m->dIDot.p, m->dIMs.p, m->dBDot.p, m->dBMs.p};
}

__device__ __forceinline__ Dv2 ld2(const double *a, int i) { return Dv2{a[2 * i], a[2 * i + 1]}; }

// fv::laplacian, all four forms (Laplacian.h:7-167) and the Vector2D theta
// forms with the SYMMETRY tangential term (Laplacian.cpp:5-120).  Same
// coefficient (gamma * dot(r, Sf)) / |r|^2 as the oracle; err bit 1 =
// unsupported boundary type, bit 2 = missing boundary/old-time input.
__global__ void kEqLaplacian(MeshDev M, int ncomp, int variant, double gamma, const double *__restrict__ gF,
                             const double *__restrict__ g0F, double theta, const int *__restrict__ bc,
                             const double *__restrict__ phiF, const double *__restrict__ phi0C,
                             const double *__restrict__ phi0F, int stride, int *cols, double *vals, int *cnt,
                             double *rhs, int *err) {
  const bool hasTheta = (variant == 0 || variant == 2);
  const bool field = (variant >= 2);
  const int N = M.nCells;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
    RowB rb[2];
    rb[0].init();
    rb[1].init();
    double src[2] = {0., 0.};
    auto addS = [&](int nb, double v) {
#pragma unroll
      for (int q = 0; q < 2; ++q)
        if (q < ncomp) rb[q].add(q * N + nb, v);
    };
    for (int k = M.iptr[c]; k < M.iptr[c + 1]; ++k) {
      const int f = M.iface[k], nb = M.inb[k];
      const double g = field ? gF[f] : gamma;
      const double coeff = g * M.idot[k] / M.ims[k];
      if (!hasTheta) {
        if (!field) {  // Laplacian.h:57-58
          addS(nb, coeff);
          addS(c, -coeff);
        } else {  // Laplacian.h:141-142
          addS(c, -coeff);
          addS(nb, coeff);
        }
        continue;
      }
      if (!phi0C) {
        atomicOr(err, 2);
        continue;
      }
      const double coeff0 = field ? g0F[f] * M.idot[k] / M.ims[k] : coeff;
      if (!field) {  // Laplacian.h:17-21
        addS(nb, theta * coeff);
        addS(c, theta * -coeff);
      } else {  // Laplacian.h:99-101, Laplacian.cpp:75-77
        addS(c, theta * -coeff);
        addS(nb, theta * coeff);
      }
      for (int q = 0; q < 2; ++q)
        if (q < ncomp) src[q] += (1. - theta) * coeff0 * (phi0C[ncomp * nb + q] - phi0C[ncomp * c + q]);
    }
    const Dv2 pc = ld2(M.cc, c);
    for (int k = M.bptr[c]; k < M.bptr[c + 1]; ++k) {
      const int f = M.bface[k];
      const int bt = bc ? bc[f] : PHASE_HIP_BC_NORMAL_GRADIENT;
      const double g = field ? gF[f] : gamma;
      const double coeff = g * M.bdot[k] / M.bms[k];
      if (bt == PHASE_HIP_BC_FIXED) {
        if (!phiF || (hasTheta && (!phi0F || !phi0C))) {
          atomicOr(err, 2);
          continue;
        }
        if (!hasTheta) {  // Laplacian.h:69-72
          addS(c, -coeff);
          for (int q = 0; q < 2; ++q) if (q < ncomp) src[q] += coeff * phiF[ncomp * f + q];
        } else {  // Laplacian.h:31-36
          const double coeff0 = field ? g0F[f] * M.bdot[k] / M.bms[k] : coeff;
          addS(c, theta * -coeff);
          for (int q = 0; q < 2; ++q) if (q < ncomp) src[q] += theta * coeff * phiF[ncomp * f + q];
          for (int q = 0; q < 2; ++q)
            if (q < ncomp) src[q] += (1. - theta) * coeff0 * (phi0F[ncomp * f + q] - phi0C[ncomp * c + q]);
        }
      } else if (bt == PHASE_HIP_BC_NORMAL_GRADIENT) {
      } else if (bt == PHASE_HIP_BC_SYMMETRY) {
        if (ncomp == 2 && hasTheta) {  // Laplacian.cpp:36-43
          if (!phi0C) {
            atomicOr(err, 2);
            continue;
          }
          const Dv2 sfo = outward(ld2(M.fc, f), ld2(M.sf, f), pc);
          const Dv2 tv{-sfo.y, sfo.x};  // Vector2D::tangentVec, Vector2D.h:28
          const double tm = sqrt(tv.x * tv.x + tv.y * tv.y);
          const Dv2 tw{tv.x / tm, tv.y / tm};
          const double coeff0 = field ? g0F[f] * M.bdot[k] / M.bms[k] : coeff;
          addS(c, theta * -coeff);
          const double tc = theta * coeff;
          // add(cell, cell, Tensor2D), VectorFiniteVolumeEquation.cpp:48-66
          const double xx = tc * (tw.x * tw.x), xy = tc * (tw.x * tw.y), yx = tc * (tw.y * tw.x),
                       yy = tc * (tw.y * tw.y);
          rb[0].add(c, xx);
          if (xy != 0.) rb[0].add(N + c, xy);
          if (yx != 0.) rb[1].add(c, yx);
          rb[1].add(N + c, yy);
          const double p0x = phi0C[2 * c], p0y = phi0C[2 * c + 1];
          const double d = p0x * tw.x + p0y * tw.y;
          const double w = (1. - theta) * coeff0;
          src[0] += w * (tw.x * d - p0x);
          src[1] += w * (tw.y * d - p0y);
        }
      } else {
        atomicOr(err, 1);
      }
    }
#pragma unroll
    for (int q = 0; q < 2; ++q)
      if (q < ncomp) {
        rb[q].store(q * N + c, stride, cols, vals, cnt, err);
        rhs[q * N + c] = src[q];
      }
  }
}

// fv::ddt(field, timeStep), TimeDerivative.h:37-48
__global__ void kEqDdt(MeshDev M, int ncomp, double dt, const double *__restrict__ phi0C, int stride, int *cols,
                       double *vals, int *cnt, double *rhs) {
  const int N = M.nCells;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
    for (int q = 0; q < ncomp; ++q) {
      const size_t row = (size_t)q * N + c;
      cols[row * stride] = (int)row;
      vals[row * stride] = M.vol[c] / dt;
      for (int j = 1; j < stride; ++j) {
        cols[row * stride + j] = -1;
        vals[row * stride + j] = 0.;
      }
      cnt[row] = 1;
      rhs[row] = -M.vol[c] * phi0C[ncomp * c + q] / dt;
    }
  }
}

// fv::ddt with a density: scalar rho (TimeDerivative.h:7-19) or a field rho with
// rho0 for the source (:21-35); mask (optional) keeps the rows of a cell
// group only (:50-62), the others stay empty
__global__ void kEqDdtRho(MeshDev M, int ncomp, double dt, double rho, const double *__restrict__ rhoC,
                          const double *__restrict__ rho0C, const double *__restrict__ phi0C,
                          const int *__restrict__ mask, int stride, int *cols, double *vals, int *cnt, double *rhs) {
  const int N = M.nCells;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
    const bool in = !mask || mask[c];
    for (int q = 0; q < ncomp; ++q) {
      const size_t row = (size_t)q * N + c;
      for (int j = 1; j < stride; ++j) {
        cols[row * stride + j] = -1;
        vals[row * stride + j] = 0.;
      }
      if (!in) {
        cols[row * stride] = -1;
        vals[row * stride] = 0.;
        cnt[row] = 0;
        rhs[row] = 0.;
        continue;
      }
      cols[row * stride] = (int)row;
      cnt[row] = 1;
      if (rhoC) {
        vals[row * stride] = rhoC[c] * M.vol[c] / dt;
        rhs[row] = -rho0C[c] * M.vol[c] * phi0C[ncomp * c + q] / dt;
      } else {
        vals[row * stride] = rho * M.vol[c] / dt;
        rhs[row] = -rho * M.vol[c] * phi0C[ncomp * c + q] / dt;
      }
    }
  }
}

__device__ __forceinline__ double upMax(double a) { return a < 0. ? 0. : a; }  // std::max(a, 0.)
__device__ __forceinline__ double upMin(double a) { return 0. < a ? 0. : a; }  // std::min(a, 0.)

// fv::div(u, phi, theta), first-order upwind, Divergence.h:7-53
__global__ void kEqDiv(MeshDev M, int ncomp, const double *__restrict__ uF, const double *__restrict__ u0F,
                       double theta, const int *__restrict__ bc, const double *__restrict__ phiF,
                       const double *__restrict__ phi0C, const double *__restrict__ phi0F, int stride, int *cols,
                       double *vals, int *cnt, double *rhs, int *err) {
  const int N = M.nCells;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
    RowB rb[2];
    rb[0].init();
    rb[1].init();
    double src[2] = {0., 0.};
    auto addS = [&](int nb, double v) {
#pragma unroll
      for (int q = 0; q < 2; ++q)
        if (q < ncomp) rb[q].add(q * N + nb, v);
    };
    const Dv2 pc = ld2(M.cc, c);
    for (int k = M.iptr[c]; k < M.iptr[c + 1]; ++k) {
      const int f = M.iface[k], nb = M.inb[k];
      const Dv2 s = outward(ld2(M.fc, f), ld2(M.sf, f), pc);
      const double flux = uF[2 * f] * s.x + uF[2 * f + 1] * s.y;
      const double flux0 = u0F[2 * f] * s.x + u0F[2 * f + 1] * s.y;
      addS(c, theta * upMax(flux));
      addS(nb, theta * upMin(flux));
      for (int q = 0; q < 2; ++q) if (q < ncomp) src[q] += (1. - theta) * upMax(flux0) * phi0C[ncomp * c + q];
      for (int q = 0; q < 2; ++q) if (q < ncomp) src[q] += (1. - theta) * upMin(flux0) * phi0C[ncomp * nb + q];
    }
    for (int k = M.bptr[c]; k < M.bptr[c + 1]; ++k) {
      const int f = M.bface[k];
      const Dv2 s = outward(ld2(M.fc, f), ld2(M.sf, f), pc);
      const double flux = uF[2 * f] * s.x + uF[2 * f + 1] * s.y;
      const double flux0 = u0F[2 * f] * s.x + u0F[2 * f + 1] * s.y;
      const int bt = bc ? bc[f] : PHASE_HIP_BC_NORMAL_GRADIENT;
      if (bt == PHASE_HIP_BC_FIXED) {
        if (!phiF || !phi0F) {
          atomicOr(err, 2);
          continue;
        }
        for (int q = 0; q < 2; ++q) if (q < ncomp) src[q] += theta * flux * phiF[ncomp * f + q];
        for (int q = 0; q < 2; ++q) if (q < ncomp) src[q] += (1. - theta) * flux0 * phi0F[ncomp * f + q];
      } else if (bt == PHASE_HIP_BC_NORMAL_GRADIENT) {
        addS(c, theta * flux);
        for (int q = 0; q < 2; ++q) if (q < ncomp) src[q] += (1. - theta) * flux0 * phi0C[ncomp * c + q];
      } else if (bt == PHASE_HIP_BC_SYMMETRY) {
      } else {
        atomicOr(err, 1);
      }
    }
#pragma unroll
    for (int q = 0; q < 2; ++q)
      if (q < ncomp) {
        rb[q].store(q * N + c, stride, cols, vals, cnt, err);
        rhs[q * N + c] = src[q];
      }
  }
}

// fv::divc(u, phi, theta), central, Divergence.h:55-103 (boundary flux0 from u,
// as the reference computes it)
__global__ void kEqDivc(MeshDev M, int ncomp, const double *__restrict__ uF, const double *__restrict__ u0F,
                        double theta, const int *__restrict__ bc, const double *__restrict__ phiF,
                        const double *__restrict__ phi0C, const double *__restrict__ phi0F, int stride, int *cols,
                        double *vals, int *cnt, double *rhs, int *err) {
  const int N = M.nCells;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
    RowB rb[2];
    rb[0].init();
    rb[1].init();
    double src[2] = {0., 0.};
    auto addS = [&](int nb, double v) {
#pragma unroll
      for (int q = 0; q < 2; ++q)
        if (q < ncomp) rb[q].add(q * N + nb, v);
    };
    const Dv2 pc = ld2(M.cc, c);
    for (int k = M.iptr[c]; k < M.iptr[c + 1]; ++k) {
      const int f = M.iface[k], nb = M.inb[k];
      const Dv2 fc = ld2(M.fc, f), pn = ld2(M.cc, nb);
      const Dv2 s = outward(fc, ld2(M.sf, f), pc);
      const double flux = theta * (uF[2 * f] * s.x + uF[2 * f + 1] * s.y);
      const double flux0 = (1. - theta) * (u0F[2 * f] * s.x + u0F[2 * f + 1] * s.y);
      const double ax = fc.x - pc.x, ay = fc.y - pc.y, bx = fc.x - pn.x, by = fc.y - pn.y;
      const double lc = sqrt(ax * ax + ay * ay), ln = sqrt(bx * bx + by * by);
      const double g = ln / (lc + ln);
      addS(c, g * flux);
      addS(nb, (1. - g) * flux);
      for (int q = 0; q < 2; ++q)
        if (q < ncomp) src[q] += flux0 * (g * phi0C[ncomp * c + q] + (1. - g) * phi0C[ncomp * nb + q]);
    }
    for (int k = M.bptr[c]; k < M.bptr[c + 1]; ++k) {
      const int f = M.bface[k];
      const Dv2 s = outward(ld2(M.fc, f), ld2(M.sf, f), pc);
      const double fl = uF[2 * f] * s.x + uF[2 * f + 1] * s.y;
      const double flux = theta * fl, flux0 = (1. - theta) * fl;
      const int bt = bc ? bc[f] : PHASE_HIP_BC_NORMAL_GRADIENT;
      if (bt == PHASE_HIP_BC_FIXED) {
        if (!phiF || !phi0F) {
          atomicOr(err, 2);
          continue;
        }
        for (int q = 0; q < 2; ++q) if (q < ncomp) src[q] += flux * phiF[ncomp * f + q];
        for (int q = 0; q < 2; ++q) if (q < ncomp) src[q] += flux0 * phi0F[ncomp * f + q];
      } else if (bt == PHASE_HIP_BC_NORMAL_GRADIENT) {
        addS(c, flux);
        for (int q = 0; q < 2; ++q) if (q < ncomp) src[q] += flux0 * phi0C[ncomp * c + q];
      } else if (bt == PHASE_HIP_BC_SYMMETRY) {
      } else {
        atomicOr(err, 1);
      }
    }
#pragma unroll
    for (int q = 0; q < 2; ++q)
      if (q < ncomp) {
        rb[q].store(q * N + c, stride, cols, vals, cnt, err);
        rhs[q * N + c] = src[q];
      }
  }
}

// fv::dive(u, phi, theta), ExplicitDivergence.h:7-51: sources only, from
// (u0, phi0) and (u1, phi1); the rows stay empty
__global__ void kEqDive(MeshDev M, int ncomp, const double *__restrict__ u0F, const double *__restrict__ u1F,
                        double theta, const int *__restrict__ bc, const double *__restrict__ phi0C,
                        const double *__restrict__ phi1C, const double *__restrict__ phi0F,
                        const double *__restrict__ phi1F, int stride, int *cols, double *vals, int *cnt,
                        double *rhs, int *err) {
  const int N = M.nCells;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
    double src[2] = {0., 0.};
    const Dv2 pc = ld2(M.cc, c);
    for (int k = M.iptr[c]; k < M.iptr[c + 1]; ++k) {
      const int f = M.iface[k], nb = M.inb[k];
      const Dv2 s = outward(ld2(M.fc, f), ld2(M.sf, f), pc);
      const double flux0 = u0F[2 * f] * s.x + u0F[2 * f + 1] * s.y;
      const double flux1 = u1F[2 * f] * s.x + u1F[2 * f + 1] * s.y;
      for (int q = 0; q < 2; ++q) if (q < ncomp) src[q] += theta * upMax(flux0) * phi0C[ncomp * c + q];
      for (int q = 0; q < 2; ++q) if (q < ncomp) src[q] += theta * upMin(flux0) * phi0C[ncomp * nb + q];
      for (int q = 0; q < 2; ++q) if (q < ncomp) src[q] += (1. - theta) * upMax(flux1) * phi1C[ncomp * c + q];
      for (int q = 0; q < 2; ++q) if (q < ncomp) src[q] += (1. - theta) * upMin(flux1) * phi1C[ncomp * nb + q];
    }
    for (int k = M.bptr[c]; k < M.bptr[c + 1]; ++k) {
      const int f = M.bface[k];
      const Dv2 s = outward(ld2(M.fc, f), ld2(M.sf, f), pc);
      const double flux0 = u0F[2 * f] * s.x + u0F[2 * f + 1] * s.y;
      const double flux1 = u1F[2 * f] * s.x + u1F[2 * f + 1] * s.y;
      const int bt = bc ? bc[f] : PHASE_HIP_BC_NORMAL_GRADIENT;
      if (bt == PHASE_HIP_BC_FIXED || bt == PHASE_HIP_BC_NORMAL_GRADIENT) {
        for (int q = 0; q < 2; ++q) if (q < ncomp) src[q] += theta * flux0 * phi0F[ncomp * f + q];
        for (int q = 0; q < 2; ++q) if (q < ncomp) src[q] += (1. - theta) * flux1 * phi1F[ncomp * f + q];
      } else if (bt == PHASE_HIP_BC_SYMMETRY) {
      } else {
        atomicOr(err, 1);
      }
    }
#pragma unroll
    for (int q = 0; q < 2; ++q)
      if (q < ncomp) {
        const int row = q * N + c;
        for (int j = 0; j < stride; ++j) {
          cols[(size_t)row * stride + j] = -1;
          vals[(size_t)row * stride + j] = 0.;
        }
        cnt[row] = 0;
        rhs[row] = src[q];
      }
  }
}

// fv::div(phiU, gradU, u), Divergence.cpp:3-47 (Vector2D): upwind with the
// upwind cell's Jacobian (xx xy yx yy) carried to the face centroid
__global__ void kEqDivGrad(MeshDev M, const double *__restrict__ phiUF, const double *__restrict__ J,
                           const int *__restrict__ bc, const double *__restrict__ uF, int stride, int *cols,
                           double *vals, int *cnt, double *rhs, int *err) {
  const int N = M.nCells;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
    RowB rb[2];
    rb[0].init();
    rb[1].init();
    double src[2] = {0., 0.};
    auto addS = [&](int nb, double v) {
      rb[0].add(nb, v);
      rb[1].add(N + nb, v);
    };
    auto jsrc = [&](int cell, double rx, double ry, double flux) {
      const double *j = J + 4 * cell;
      const double vx = j[0] * rx + j[1] * ry, vy = j[2] * rx + j[3] * ry;
      src[0] += flux * vx;
      src[1] += flux * vy;
    };
    const Dv2 pc = ld2(M.cc, c);
    for (int k = M.iptr[c]; k < M.iptr[c + 1]; ++k) {
      const int f = M.iface[k], nb = M.inb[k];
      const Dv2 fc = ld2(M.fc, f);
      const Dv2 s = outward(fc, ld2(M.sf, f), pc);
      const double flux = phiUF[2 * f] * s.x + phiUF[2 * f + 1] * s.y;
      if (flux > 0.) {
        addS(c, flux);
        jsrc(c, fc.x - pc.x, fc.y - pc.y, flux);
      } else {
        const Dv2 pn = ld2(M.cc, nb);
        addS(nb, flux);
        jsrc(nb, fc.x - pn.x, fc.y - pn.y, flux);
      }
    }
    for (int k = M.bptr[c]; k < M.bptr[c + 1]; ++k) {
      const int f = M.bface[k];
      const Dv2 fc = ld2(M.fc, f);
      const Dv2 s = outward(fc, ld2(M.sf, f), pc);
      const double flux = phiUF[2 * f] * s.x + phiUF[2 * f + 1] * s.y;
      const int bt = bc ? bc[f] : PHASE_HIP_BC_NORMAL_GRADIENT;
      if (bt == PHASE_HIP_BC_FIXED) {
        if (!uF) {
          atomicOr(err, 2);
          continue;
        }
        src[0] += flux * uF[2 * f];
        src[1] += flux * uF[2 * f + 1];
      } else if (bt == PHASE_HIP_BC_NORMAL_GRADIENT) {
        addS(c, flux);
        jsrc(c, fc.x - pc.x, fc.y - pc.y, flux);
      } else if (bt == PHASE_HIP_BC_SYMMETRY) {
      } else {
        atomicOr(err, 1);
      }
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      rb[q].store(q * N + c, stride, cols, vals, cnt, err);
      rhs[q * N + c] = src[q];
    }
  }
}

// VectorFiniteVolumeField::faceToCell, plain (VectorFiniteVolumeField.cpp:5-26)
// or weighted (:28-51): sum of pointwise(v_f, |S_f|) [/ faceWeight_f] over the
// cell's faces, componentwise over sum |S_f| [times cellWeight]
__global__ void kFaceToCell(MeshDev M, const double *__restrict__ fv, const double *__restrict__ cw,
                            const double *__restrict__ fw, double *__restrict__ out) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < M.nCells; c += gridDim.x * blockDim.x) {
    double sx = 0., sy = 0., tx = 0., ty = 0.;
    auto add = [&](int f) {
      const Dv2 n = ld2(M.sf, f);
      const double ax = fabs(n.x), ay = fabs(n.y);
      double px = fv[2 * f] * ax, py = fv[2 * f + 1] * ay;
      if (fw) {
        px /= fw[f];
        py /= fw[f];
      }
      tx += px;
      ty += py;
      sx += ax;
      sy += ay;
    };
    for (int k = M.iptr[c]; k < M.iptr[c + 1]; ++k) add(M.iface[k]);
    for (int k = M.bptr[c]; k < M.bptr[c + 1]; ++k) add(M.bface[k]);
    double vx = tx / sx, vy = ty / sy;
    if (cw) {
      vx = cw[c] * vx;
      vy = cw[c] * vy;
    }
    out[2 * c] = vx;
    out[2 * c + 1] = vy;
  }
}

// ---------------------------------------------------------------- VOF (CICSAM)
// std::min / std::max as the reference calls them -- min(a, b) = (b < a) ? b : a,
// max(a, b) = (a < b) ? b : a -- so a NaN first argument passes through
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
// clamp(v, lo, hi) = max(min(v, hi), lo), Math/Algorithm.h:11-13
__device__ __forceinline__ double sclamp(double v, double lo, double hi) { return smax(smin(v, hi), lo); }
// cicsam::hc and cicsam::uq, Cicsam.cpp:5-18
__device__ __forceinline__ double cicsamHc(double g, double co) { return g >= 0 && g <= 1 ? smin(1., g / co) : g; }
__device__ __forceinline__ double cicsamUq(double g, double co) {
  return g >= 0 && g <= 1 ? smin((8. * co * g + (1. - co) * (6. * g + 3.)) / 8., cicsamHc(g, co)) : g;
}

// cicsam::faceInterpolationWeights, Cicsam.cpp:20-66 (k = 1), one face per
// thread; boundary faces 0.  cos(2 theta_F) with theta_F = acos(c) is
// evaluated as 2 c^2 - 1 (c > 1 from rounding stays NaN, as acos makes it).
__global__ void kCicsamWeights(MeshDev M, int nFaces, const int *__restrict__ fl, const int *__restrict__ fr,
                               const double *__restrict__ uF, const double *__restrict__ gC,
                               const double *__restrict__ gradC, double dt, double *__restrict__ beta) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nFaces; f += gridDim.x * blockDim.x) {
    const int l = fl[f], r = fr[f];
    if (r < 0) {
      beta[f] = 0.;
      continue;
    }
    const Dv2 uf = ld2(uF, f);
    const Dv2 s = outward(ld2(M.fc, f), ld2(M.sf, f), ld2(M.cc, l));
    const double flux = uf.x * s.x + uf.y * s.y;
    const int d = flux > 0. ? l : r, a = flux <= 0. ? l : r;
    const Dv2 cd = ld2(M.cc, d), ca = ld2(M.cc, a);
    const Dv2 rc{ca.x - cd.x, ca.y - cd.y};
    const Dv2 gd = ld2(gradC, d);
    const double gD = sclamp(gC[d], 0., 1.), gA = sclamp(gC[a], 0., 1.);
    const double gU = sclamp(gA - 2. * (rc.x * gd.x + rc.y * gd.y), 0., 1.);
    double gDT = (gD - gU) / (gA - gU);
    if (!isfinite(gDT)) gDT = 0.;
    // donor Courant number over all of the donor's faces
    const double vd = M.vol[d];
    double coD = 0.;
    for (int k = M.iptr[d]; k < M.iptr[d + 1]; ++k) {
      const int g = M.iface[k];
      const Dv2 n = outward(ld2(M.fc, g), ld2(M.sf, g), cd);
      const Dv2 ug = ld2(uF, g);
      coD += smax((ug.x * n.x + ug.y * n.y) / vd * dt, 0.);
    }
    for (int k = M.bptr[d]; k < M.bptr[d + 1]; ++k) {
      const int g = M.bface[k];
      const Dv2 n = outward(ld2(M.fc, g), ld2(M.sf, g), cd);
      const Dv2 ug = ld2(uF, g);
      coD += smax((ug.x * n.x + ug.y * n.y) / vd * dt, 0.);
    }
    const double mg = sqrt(gd.x * gd.x + gd.y * gd.y), mr = sqrt(rc.x * rc.x + rc.y * rc.y);
    const double c = fabs((gd.x / mg) * (rc.x / mr) + (gd.y / mg) * (rc.y / mr));
    const double cos2 = c > 1. ? __builtin_nan("") : 2. * c * c - 1.;
    const double psiF = smin(1. * (cos2 + 1.) / 2., 1.);
    const double gFT = psiF * cicsamHc(gDT, coD) + (1. - psiF) * cicsamUq(gDT, coD);
    const double b = (gFT - gDT) / (1. - gDT);
    beta[f] = isfinite(b) ? smax(smin(1., b), 0.) : 0.;
  }
}

// cicsam::div(u, gamma, beta, theta), Cicsam.cpp:87-138: per link the donor and
// acceptor coefficients theta (1 - b) F and theta b F in that order, the
// explicit part as a source, boundary links by type (err bit 1: other type)
__global__ void kEqCicsamDiv(MeshDev M, const double *__restrict__ uF, const double *__restrict__ beta, double theta,
                             const int *__restrict__ bc, const double *__restrict__ gF,
                             const double *__restrict__ g0C, const double *__restrict__ g0F, int stride, int *cols,
                             double *vals, int *cnt, double *rhs, int *err) {
  const int N = M.nCells;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
    RowB rb;
    rb.init();
    double src = 0.;
    const Dv2 pc = ld2(M.cc, c);
    for (int k = M.iptr[c]; k < M.iptr[c + 1]; ++k) {
      const int f = M.iface[k], nb = M.inb[k];
      const Dv2 s = outward(ld2(M.fc, f), ld2(M.sf, f), pc);
      const double flux = uF[2 * f] * s.x + uF[2 * f + 1] * s.y;
      const int d = flux > 0. ? c : nb, a = flux <= 0. ? c : nb;
      const double b = beta[f];
      rb.add(d, theta * (1. - b) * flux);
      rb.add(a, theta * b * flux);
      const double gFace = (1. - b) * g0C[d] + b * g0C[a];
      src += (1. - theta) * flux * gFace;
    }
    for (int k = M.bptr[c]; k < M.bptr[c + 1]; ++k) {
      const int f = M.bface[k];
      const Dv2 s = outward(ld2(M.fc, f), ld2(M.sf, f), pc);
      const double flux = uF[2 * f] * s.x + uF[2 * f + 1] * s.y;
      const int bt = bc ? bc[f] : PHASE_HIP_BC_NORMAL_GRADIENT;
      if (bt == PHASE_HIP_BC_FIXED) {
        if (!gF || !g0F) {
          atomicOr(err, 2);
          continue;
        }
        src += theta * flux * gF[f];
        src += (1. - theta) * flux * g0F[f];
      } else if (bt == PHASE_HIP_BC_NORMAL_GRADIENT) {
        rb.add(c, theta * flux);
        src += (1. - theta) * flux * g0C[c];
      } else if (bt == PHASE_HIP_BC_SYMMETRY) {
      } else {
        atomicOr(err, 1);
      }
    }
    rb.store(c, stride, cols, vals, cnt, err);
    rhs[c] = src;
  }
}

// cicsam::computeMomentumFlux, Cicsam.cpp:68-85 (rhoU per face, interleaved)
__global__ void kCicsamMomentumFlux(MeshDev M, int nFaces, const int *__restrict__ fl, const int *__restrict__ fr,
                                    double rho1, double rho2, const double *__restrict__ uF,
                                    const double *__restrict__ gC, const double *__restrict__ gF,
                                    const double *__restrict__ beta, double *__restrict__ rhoU) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nFaces; f += gridDim.x * blockDim.x) {
    const int l = fl[f], r = fr[f];
    const Dv2 uf = ld2(uF, f);
    double g;
    if (r >= 0) {
      const Dv2 s = outward(ld2(M.fc, f), ld2(M.sf, f), ld2(M.cc, l));
      const double flux = uf.x * s.x + uf.y * s.y;
      const int d = flux > 0. ? l : r, a = flux <= 0. ? l : r;
      const double b = beta[f];
      g = (1. - b) * gC[d] + b * gC[a];
    } else {
      g = gF[f];
    }
    const double rho = rho1 + sclamp(g, 0., 1.) * (rho2 - rho1);
    rhoU[2 * f] = rho * uf.x;
    rhoU[2 * f + 1] = rho * uf.y;
  }
}

// FractionalStepMultiphase::updateProperties, FractionalStepMultiphase.cpp:175-214:
// density, and dynamic viscosity from the kinematic viscosities, for n values
__global__ void kVofProperties(int n, double rho1, double rho2, double mu1, double mu2, const double *__restrict__ g,
                               double *__restrict__ rho, double *__restrict__ mu) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double gi = sclamp(g[i], 0., 1.);
    const double r = rho1 + gi * (rho2 - rho1);
    rho[i] = r;
    mu[i] = r / (rho1 / mu1 + gi * (rho2 / mu2 - rho1 / mu1));
  }
}

// lhs (+|-)= rhs, CrsEquation.cpp:185-275
__global__ void kEqMerge(int rank, int stride, int sign, int *lc, double *lv, int *ln, double *lr,
                         const int *__restrict__ rc, const double *__restrict__ rv, const int *__restrict__ rn,
                         const double *__restrict__ rr, int *err) {
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < rank; row += gridDim.x * blockDim.x) {
    RowB rb;
    rb.init();
    const size_t o = (size_t)row * stride;
    // the left row's non-zeros keep their order: distinct columns, so add() appends
    for (int j = 0; j < ln[row]; ++j)
      if (lv[o + j] != 0. && lc[o + j] >= 0) rb.add(lc[o + j], lv[o + j]);
    for (int j = 0; j < rn[row]; ++j) {
      const double v = rv[o + j];
      if (v == 0. || rc[o + j] < 0) continue;
      rb.add(rc[o + j], sign > 0 ? v : -v);
    }
    rb.store(row, stride, lc, lv, ln, err);
    lr[row] = sign > 0 ? lr[row] + rr[row] : lr[row] - rr[row];
  }
}

// CrsEquation::operator*=(Scalar), CrsEquation.cpp:287-291
__global__ void kEqScale(int rank, int stride, double s, double *vals, double *rhs) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)rank * stride;
       i += (size_t)gridDim.x * blockDim.x)
    vals[i] *= s;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < rank; i += gridDim.x * blockDim.x) rhs[i] *= s;
}

// operator*(ScalarFiniteVolumeField, equation): every slot and the rhs of row
// r times field[r mod nCells] (scaleRow, CrsEquation.cpp:153-159)
__global__ void kEqScaleCells(int rank, int nCells, int stride, const double *__restrict__ field, double *vals,
                              double *rhs) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)rank * stride;
       i += (size_t)gridDim.x * blockDim.x)
    vals[i] *= field[(i / stride) % nCells];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < rank; i += gridDim.x * blockDim.x)
    rhs[i] *= field[i % nCells];
}

// remove(cell): the masked cells' rows keep their values but lose their
// columns (-1) and rhs
__global__ void kEqRemoveCells(int rank, int nCells, int stride, const int *__restrict__ mask, int *cols,
                               double *rhs) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rank; r += gridDim.x * blockDim.x) {
    if (!mask[r % nCells]) continue;
    for (int j = 0; j < stride; ++j) cols[(size_t)r * stride + j] = -1;
    rhs[r] = 0.;
  }
}

// rhs -= v (operator-=(Vector) / ==(Vector), CrsEquation.cpp:282-285, 308-311);
// v == nullptr: rhs -= s (operator==(Scalar), :299-303)
__global__ void kEqSubRhs(int rank, const double *__restrict__ v, double s, double *rhs) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < rank; i += gridDim.x * blockDim.x)
    rhs[i] -= v ? v[i] : s;
}

// implicit under-relaxation: a_P /= alpha; rhs -= (1 - alpha) * a_P * phi
// (ScalarFiniteVolumeEquation.cpp:45-55, VectorFiniteVolumeEquation.cpp:129-146)
__global__ void kEqRelax(int rank, int stride, double alpha, const double *__restrict__ phi, const int *cols,
                         double *vals, const int *cnt, double *rhs) {
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < rank; row += gridDim.x * blockDim.x) {
    const size_t o = (size_t)row * stride;
    for (int j = 0; j < cnt[row]; ++j)
      if (cols[o + j] == row) {
        vals[o + j] /= alpha;
        rhs[row] -= (1. - alpha) * vals[o + j] * phi[row];
        break;
      }
  }
}

// src::div(u), Source.cpp:5-21: sum over the cell's faces of dot(u_f, Sf_out)
__global__ void kSrcDiv(MeshDev M, const double *__restrict__ uF, double *out) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < M.nCells; c += gridDim.x * blockDim.x) {
    const Dv2 pc = ld2(M.cc, c);
    double d = 0.;
    for (int k = M.iptr[c]; k < M.iptr[c + 1]; ++k) {
      const int f = M.iface[k];
      const Dv2 s = outward(ld2(M.fc, f), ld2(M.sf, f), pc);
      d += uF[2 * f] * s.x + uF[2 * f + 1] * s.y;
    }
    for (int k = M.bptr[c]; k < M.bptr[c + 1]; ++k) {
      const int f = M.bface[k];
      const Dv2 s = outward(ld2(M.fc, f), ld2(M.sf, f), pc);
      d += uF[2 * f] * s.x + uF[2 * f + 1] * s.y;
    }
    out[c] = d;
  }
}

// src::src(field), Source.cpp:77-95: phi * V per row (block layout)
__global__ void kSrcSrc(MeshDev M, int ncomp, const double *__restrict__ phiC, double *out) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < M.nCells; c += gridDim.x * blockDim.x)
    for (int q = 0; q < ncomp; ++q) out[(size_t)q * M.nCells + c] = phiC[ncomp * c + q] * M.vol[c];
}

// ScalarGradient::computeFaces, ScalarGradient.cpp:34-52
__global__ void kGradFaces(int nFaces, const int *__restrict__ fl, const int *__restrict__ fr,
                           const double *__restrict__ cc, const double *__restrict__ fc,
                           const double *__restrict__ phiC, const double *__restrict__ phiF, double *gF) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nFaces; f += gridDim.x * blockDim.x) {
    const int l = fl[f], r = fr[f];
    double dx, dy, dphi;
    if (r >= 0) {
      dx = cc[2 * r] - cc[2 * l];
      dy = cc[2 * r + 1] - cc[2 * l + 1];
      dphi = phiC[r] - phiC[l];
    } else {
      dx = fc[2 * f] - cc[2 * l];
      dy = fc[2 * f + 1] - cc[2 * l + 1];
      dphi = phiF[f] - phiC[l];
    }
    const double ms = dx * dx + dy * dy;
    gF[2 * f] = dx * dphi / ms;
    gF[2 * f + 1] = dy * dphi / ms;
  }
}

// ScalarGradient::compute(FACE_TO_CELL), ScalarGradient.cpp:54-74:
// componentwise sum_f g_f |Sf_out| / sum_f |Sf_out| over the cell's links
__global__ void kGradCells(MeshDev M, const double *__restrict__ gF, double *gC) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < M.nCells; c += gridDim.x * blockDim.x) {
    const Dv2 pc = ld2(M.cc, c);
    double sx = 0., sy = 0., tx = 0., ty = 0.;
    auto acc = [&](int f) {
      const Dv2 n = outward(ld2(M.fc, f), ld2(M.sf, f), pc);
      const double ax = fabs(n.x), ay = fabs(n.y);
      tx += gF[2 * f] * ax;
      ty += gF[2 * f + 1] * ay;
      sx += ax;
      sy += ay;
    };
    for (int k = M.iptr[c]; k < M.iptr[c + 1]; ++k) acc(M.iface[k]);
    for (int k = M.bptr[c]; k < M.bptr[c + 1]; ++k) acc(M.bface[k]);
    gC[2 * c] = tx / sx;
    gC[2 * c + 1] = ty / sy;
  }
}

// interpolateFaces(DISTANCE) + setBoundaryFaces, FiniteVolumeField.tpp:129-182
// and VectorFiniteVolumeField.cpp:140-160; FIXED faces keep their value.
__global__ void kInterpFaces(int nFaces, int ncomp, const int *__restrict__ fl, const int *__restrict__ fr,
                             const double *__restrict__ cc, const double *__restrict__ fc,
                             const double *__restrict__ sf, const int *__restrict__ bc,
                             const double *__restrict__ phiC, double *phiF) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nFaces; f += gridDim.x * blockDim.x) {
    const int l = fl[f], r = fr[f];
    if (r >= 0) {
      const double ax = fc[2 * f] - cc[2 * r], ay = fc[2 * f + 1] - cc[2 * r + 1];
      const double bx = fc[2 * f] - cc[2 * l], by = fc[2 * f + 1] - cc[2 * l + 1];
      const double l1 = sqrt(ax * ax + ay * ay), l2 = sqrt(bx * bx + by * by);
      const double g = l1 / (l1 + l2);  // Face::distanceWeight, Face.cpp:60-64
      for (int q = 0; q < ncomp; ++q)
        phiF[ncomp * f + q] = g * phiC[ncomp * l + q] + (1. - g) * phiC[ncomp * r + q];
      continue;
    }
    const int bt = bc ? bc[f] : PHASE_HIP_BC_NORMAL_GRADIENT;
    if (bt == PHASE_HIP_BC_FIXED) continue;
    if (bt == PHASE_HIP_BC_SYMMETRY && ncomp == 2) {
      const double nx = sf[2 * f], ny = sf[2 * f + 1];
      const double px = phiC[2 * l], py = phiC[2 * l + 1];
      const double d = px * nx + py * ny;
      const double ms = nx * nx + ny * ny;
      phiF[2 * f] = px - nx * d / ms;
      phiF[2 * f + 1] = py - ny * d / ms;
    } else {
      for (int q = 0; q < ncomp; ++q) phiF[ncomp * f + q] = phiC[ncomp * l + q];
    }
  }
}

inline int g256(long n) { return (int)std::max<long>(1, std::min<long>((n + 255) / 256, 8192)); }

template <class F>
int guardedEq(phase_hip_equation *e, F &&f) {
  if (!e) return PHASE_HIP_EINVAL;
  try {
    f();
    return PHASE_HIP_OK;
  } catch (const HipError &x) {
    e->err = x.msg;
    return x.code;
  } catch (const std::exception &x) {
    e->err = x.what();
    return PHASE_HIP_EINVAL;
  }
}

template <class F>
int guardedM(phase_hip_mesh *m, F &&f) {
  if (!m) return PHASE_HIP_EINVAL;
  try {
    f();
    return PHASE_HIP_OK;
  } catch (const HipError &x) {
    m->err = x.msg;
    return x.code;
  } catch (const std::exception &x) {
    m->err = x.what();
    return PHASE_HIP_EINVAL;
  }
}

// Device error bits of the last operator: raised as the reference's exceptions.
void checkEq(phase_hip_equation *e, const char *op) {
  int v = 0;
  PH_CHECK(hipMemcpyAsync(&v, e->dErr.p, sizeof(int), hipMemcpyDeviceToHost, e->mesh->stream));
  PH_CHECK(hipStreamSynchronize(e->mesh->stream));
  if (!v) return;
  PH_CHECK(hipMemsetAsync(e->dErr.p, 0, sizeof(int), e->mesh->stream));
  if (v & 1) fail(PHASE_HIP_EBOUNDARY, std::string(op) + ": unrecognized or unspecified boundary type.");
  if (v & 2) fail(PHASE_HIP_EINVAL, std::string(op) + ": a boundary or old-time input is missing");
  fail(PHASE_HIP_EINVAL, std::string(op) + ": a row needs more slots than the equation holds");
}

void fresh(phase_hip_equation *e) { e->compact = false; }

}  // namespace

namespace phase {

// Solver side of FiniteVolumeEquation<T>::solve (FiniteVolumeEquation.tpp:64-86):
// A from the equation's rows, b = -rhs.  Defined in solver.hip.  adjPtr /
// adjNb (host, may be NULL): the rows' structural neighbours (for scalar mesh
// equations the face neighbours), the largest pattern the rows can take.
void solverSetFromRows(phase_hip_solver *s, int rank, int stride, const int *dCols, const double *dVals,
                       const double *dRhs, hipStream_t producer, const double *dCellXY, const int *adjPtr,
                       const int *adjNb);

}  // namespace phase

extern "C" {

const char *phase_hip_equation_last_error(const phase_hip_equation *e) { return e ? e->err.c_str() : "null equation"; }

int phase_hip_equation_create(phase_hip_equation **out, phase_hip_mesh *m, int ncomp) {
  if (!out || !m || (ncomp != 1 && ncomp != 2)) return PHASE_HIP_EINVAL;
  *out = nullptr;
  phase_hip_equation *e = new phase_hip_equation();
  e->mesh = m;
  e->ncomp = ncomp;
  e->rank = ncomp * m->nCells;
  // diagonal + neighbours (+ one cross-component column for Vector2D)
  e->stride = std::max(5, 1 + m->maxDeg + (ncomp == 2 ? 1 : 0));
  const int rc = guardedEq(e, [&] {
    if (e->stride > kRowCap) fail(PHASE_HIP_EINVAL, "a cell has more faces than an equation row holds");
    PH_CHECK(hipSetDevice(m->device));
    const size_t slots = (size_t)e->rank * e->stride;
    e->cols.alloc(slots);
    e->vals.alloc(slots);
    e->cnt.alloc(e->rank);
    e->rhs.alloc(e->rank);
    e->dErr.alloc(1);
    e->dErr.zero(m->stream);
    // CrsEquation(N, 5): every slot empty
    PH_CHECK(hipMemsetAsync(e->cols.p, 0xff, slots * sizeof(int), m->stream));
    e->vals.zero(m->stream);
    e->cnt.zero(m->stream);
    e->rhs.zero(m->stream);
    PH_CHECK(hipStreamSynchronize(m->stream));
  });
  if (rc != PHASE_HIP_OK) {
    delete e;
    return rc;
  }
  *out = e;
  return PHASE_HIP_OK;
}

int phase_hip_equation_destroy(phase_hip_equation *e) {
  delete e;
  return PHASE_HIP_OK;
}

int phase_hip_equation_rank(const phase_hip_equation *e) { return e ? e->rank : 0; }

int phase_hip_eq_laplacian(phase_hip_equation *e, int variant, double gamma, const double *gammaFace,
                           const double *gamma0Face, double theta, const int *bcType, const double *phiFace,
                           const double *phi0Cell, const double *phi0Face) {
  return guardedEq(e, [&] {
    if (variant < 0 || variant > 3) fail(PHASE_HIP_EINVAL, "laplacian variant must be 0..3");
    if (variant >= 2 && (!gammaFace || (variant == 2 && !gamma0Face)))
      fail(PHASE_HIP_EINVAL, "face-field laplacian needs gammaFace (and gamma0Face with theta)");
    phase_hip_mesh *m = e->mesh;
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kEqLaplacian, dim3(g256(m->nCells)), dim3(256), 0, m->stream, meshDev(m), e->ncomp,
                       variant, gamma, gammaFace, gamma0Face, theta, bcType, phiFace, phi0Cell, phi0Face, e->stride,
                       e->cols.p, e->vals.p, e->cnt.p, e->rhs.p, e->dErr.p);
    PH_CHECK(hipGetLastError());
    fresh(e);
    checkEq(e, "fv::laplacian");
  });
}

int phase_hip_eq_ddt(phase_hip_equation *e, double timeStep, const double *phi0Cell) {
  return guardedEq(e, [&] {
    if (!phi0Cell) fail(PHASE_HIP_EINVAL, "fv::ddt needs the old-time field");
    phase_hip_mesh *m = e->mesh;
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kEqDdt, dim3(g256(m->nCells)), dim3(256), 0, m->stream, meshDev(m), e->ncomp, timeStep,
                       phi0Cell, e->stride, e->cols.p, e->vals.p, e->cnt.p, e->rhs.p);
    PH_CHECK(hipGetLastError());
    fresh(e);
  });
}

int phase_hip_eq_ddt_rho(phase_hip_equation *e, double timeStep, double rho, const double *rhoCell,
                         const double *rho0Cell, const double *phi0Cell, const int *cellMask) {
  return guardedEq(e, [&] {
    if (!phi0Cell) fail(PHASE_HIP_EINVAL, "fv::ddt needs the old-time field");
    if (!rhoCell != !rho0Cell) fail(PHASE_HIP_EINVAL, "a density field needs its old-time values too");
    phase_hip_mesh *m = e->mesh;
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kEqDdtRho, dim3(g256(m->nCells)), dim3(256), 0, m->stream, meshDev(m), e->ncomp, timeStep,
                       rho, rhoCell, rho0Cell, phi0Cell, cellMask, e->stride, e->cols.p, e->vals.p, e->cnt.p,
                       e->rhs.p);
    PH_CHECK(hipGetLastError());
    fresh(e);
  });
}

int phase_hip_eq_div(phase_hip_equation *e, const double *uFace, const double *u0Face, double theta,
                     const int *bcType, const double *phiFace, const double *phi0Cell, const double *phi0Face) {
  return guardedEq(e, [&] {
    if (!uFace || !u0Face || !phi0Cell) fail(PHASE_HIP_EINVAL, "fv::div needs u, u0 (faces) and phi0 (cells)");
    phase_hip_mesh *m = e->mesh;
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kEqDiv, dim3(g256(m->nCells)), dim3(256), 0, m->stream, meshDev(m), e->ncomp, uFace, u0Face,
                       theta, bcType, phiFace, phi0Cell, phi0Face, e->stride, e->cols.p, e->vals.p, e->cnt.p,
                       e->rhs.p, e->dErr.p);
    PH_CHECK(hipGetLastError());
    fresh(e);
    checkEq(e, "fv::div");
  });
}

int phase_hip_eq_divc(phase_hip_equation *e, const double *uFace, const double *u0Face, double theta,
                      const int *bcType, const double *phiFace, const double *phi0Cell, const double *phi0Face) {
  return guardedEq(e, [&] {
    if (!uFace || !u0Face || !phi0Cell) fail(PHASE_HIP_EINVAL, "fv::divc needs u, u0 (faces) and phi0 (cells)");
    phase_hip_mesh *m = e->mesh;
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kEqDivc, dim3(g256(m->nCells)), dim3(256), 0, m->stream, meshDev(m), e->ncomp, uFace, u0Face,
                       theta, bcType, phiFace, phi0Cell, phi0Face, e->stride, e->cols.p, e->vals.p, e->cnt.p,
                       e->rhs.p, e->dErr.p);
    PH_CHECK(hipGetLastError());
    fresh(e);
    checkEq(e, "fv::divc");
  });
}

int phase_hip_eq_dive(phase_hip_equation *e, const double *u0Face, const double *u1Face, double theta,
                      const int *bcType, const double *phi0Cell, const double *phi1Cell, const double *phi0Face,
                      const double *phi1Face) {
  return guardedEq(e, [&] {
    if (!u0Face || !u1Face || !phi0Cell || !phi1Cell || !phi0Face || !phi1Face)
      fail(PHASE_HIP_EINVAL, "fv::dive needs u0, u1, phi0, phi1 (cells and faces)");
    phase_hip_mesh *m = e->mesh;
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kEqDive, dim3(g256(m->nCells)), dim3(256), 0, m->stream, meshDev(m), e->ncomp, u0Face, u1Face,
                       theta, bcType, phi0Cell, phi1Cell, phi0Face, phi1Face, e->stride, e->cols.p, e->vals.p,
                       e->cnt.p, e->rhs.p, e->dErr.p);
    PH_CHECK(hipGetLastError());
    fresh(e);
    checkEq(e, "fv::dive");
  });
}

int phase_hip_eq_div_grad(phase_hip_equation *e, const double *phiUFace, const double *gradUCell, const int *bcType,
                          const double *uFace) {
  return guardedEq(e, [&] {
    if (e->ncomp != 2) fail(PHASE_HIP_EINVAL, "fv::div(phiU, gradU, u) builds a Vector2D equation");
    if (!phiUFace || !gradUCell) fail(PHASE_HIP_EINVAL, "fv::div needs phiU (faces) and gradU (cells)");
    phase_hip_mesh *m = e->mesh;
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kEqDivGrad, dim3(g256(m->nCells)), dim3(256), 0, m->stream, meshDev(m), phiUFace, gradUCell,
                       bcType, uFace, e->stride, e->cols.p, e->vals.p, e->cnt.p, e->rhs.p, e->dErr.p);
    PH_CHECK(hipGetLastError());
    fresh(e);
    checkEq(e, "fv::div");
  });
}

int phase_hip_eq_cicsam_div(phase_hip_equation *e, const double *uFace, const double *beta, double theta,
                            const int *bcType, const double *gammaFace, const double *gamma0Cell,
                            const double *gamma0Face) {
  return guardedEq(e, [&] {
    if (e->ncomp != 1) fail(PHASE_HIP_EINVAL, "cicsam::div builds a scalar equation");
    if (!uFace || !beta || !gamma0Cell) fail(PHASE_HIP_EINVAL, "cicsam::div needs u, beta (faces) and gamma0 (cells)");
    phase_hip_mesh *m = e->mesh;
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kEqCicsamDiv, dim3(g256(m->nCells)), dim3(256), 0, m->stream, meshDev(m), uFace, beta, theta,
                       bcType, gammaFace, gamma0Cell, gamma0Face, e->stride, e->cols.p, e->vals.p, e->cnt.p,
                       e->rhs.p, e->dErr.p);
    PH_CHECK(hipGetLastError());
    fresh(e);
    checkEq(e, "cicsam::div");
  });
}

int phase_hip_eq_add(phase_hip_equation *e, const phase_hip_equation *o, int sign) {
  return guardedEq(e, [&] {
    if (!o || o->rank != e->rank || o->stride != e->stride || o->mesh != e->mesh)
      fail(PHASE_HIP_EINVAL, "operands of += / -= must be equations of the same mesh and rank");
    phase_hip_mesh *m = e->mesh;
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kEqMerge, dim3(g256(e->rank)), dim3(256), 0, m->stream, e->rank, e->stride, sign >= 0 ? 1 : -1,
                       e->cols.p, e->vals.p, e->cnt.p, e->rhs.p, o->cols.p, o->vals.p, o->cnt.p, o->rhs.p, e->dErr.p);
    PH_CHECK(hipGetLastError());
    e->compact = true;
    checkEq(e, "CrsEquation::operator+=");
  });
}

int phase_hip_eq_copy(phase_hip_equation *dst, const phase_hip_equation *src) {
  return guardedEq(dst, [&] {
    if (!src || src->rank != dst->rank || src->stride != dst->stride)
      fail(PHASE_HIP_EINVAL, "copy needs equations of the same rank");
    hipStream_t st = dst->mesh->stream;
    const size_t slots = (size_t)dst->rank * dst->stride;
    PH_CHECK(hipMemcpyAsync(dst->cols.p, src->cols.p, slots * sizeof(int), hipMemcpyDeviceToDevice, st));
    PH_CHECK(hipMemcpyAsync(dst->vals.p, src->vals.p, slots * sizeof(double), hipMemcpyDeviceToDevice, st));
    PH_CHECK(hipMemcpyAsync(dst->cnt.p, src->cnt.p, dst->rank * sizeof(int), hipMemcpyDeviceToDevice, st));
    PH_CHECK(hipMemcpyAsync(dst->rhs.p, src->rhs.p, dst->rank * sizeof(double), hipMemcpyDeviceToDevice, st));
    dst->compact = src->compact;
  });
}

int phase_hip_eq_scale(phase_hip_equation *e, double s) {
  return guardedEq(e, [&] {
    hipLaunchKernelGGL(kEqScale, dim3(g256((long)e->rank * e->stride)), dim3(256), 0, e->mesh->stream, e->rank,
                       e->stride, s, e->vals.p, e->rhs.p);
    PH_CHECK(hipGetLastError());
  });
}

int phase_hip_eq_scale_cells(phase_hip_equation *e, const double *field) {
  return guardedEq(e, [&] {
    if (!field) fail(PHASE_HIP_EINVAL, "null field");
    hipLaunchKernelGGL(kEqScaleCells, dim3(g256((long)e->rank * e->stride)), dim3(256), 0, e->mesh->stream, e->rank,
                       e->mesh->nCells, e->stride, field, e->vals.p, e->rhs.p);
    PH_CHECK(hipGetLastError());
  });
}

int phase_hip_eq_remove_cells(phase_hip_equation *e, const int *cellMask) {
  return guardedEq(e, [&] {
    if (!cellMask) fail(PHASE_HIP_EINVAL, "null mask");
    hipLaunchKernelGGL(kEqRemoveCells, dim3(g256(e->rank)), dim3(256), 0, e->mesh->stream, e->rank,
                       e->mesh->nCells, e->stride, cellMask, e->cols.p, e->rhs.p);
    PH_CHECK(hipGetLastError());
  });
}

int phase_hip_eq_sub_vector(phase_hip_equation *e, const double *v) {
  return guardedEq(e, [&] {
    if (!v) fail(PHASE_HIP_EINVAL, "null vector");
    hipLaunchKernelGGL(kEqSubRhs, dim3(g256(e->rank)), dim3(256), 0, e->mesh->stream, e->rank, v, 0., e->rhs.p);
    PH_CHECK(hipGetLastError());
  });
}

int phase_hip_eq_eq_scalar(phase_hip_equation *e, double s) {
  return guardedEq(e, [&] {
    if (s == 0.) return;  // CrsEquation.cpp:299-303
    hipLaunchKernelGGL(kEqSubRhs, dim3(g256(e->rank)), dim3(256), 0, e->mesh->stream, e->rank, (const double *)nullptr,
                       s, e->rhs.p);
    PH_CHECK(hipGetLastError());
  });
}

int phase_hip_eq_relax(phase_hip_equation *e, double alpha, const double *phi) {
  return guardedEq(e, [&] {
    if (!phi || !(alpha > 0.)) fail(PHASE_HIP_EINVAL, "relax needs alpha > 0 and the current field");
    hipLaunchKernelGGL(kEqRelax, dim3(g256(e->rank)), dim3(256), 0, e->mesh->stream, e->rank, e->stride, alpha, phi,
                       e->cols.p, e->vals.p, e->cnt.p, e->rhs.p);
    PH_CHECK(hipGetLastError());
  });
}

int phase_hip_eq_export(phase_hip_equation *e, int *nslots, int *rowPtr, int *colInd, double *vals, double *rhs) {
  return guardedEq(e, [&] {
    if (!nslots) fail(PHASE_HIP_EINVAL, "nslots is required");
    hipStream_t st = e->mesh->stream;
    const size_t slots = (size_t)e->rank * e->stride;
    std::vector<int> c(slots), n(e->rank);
    std::vector<double> v(slots), r(e->rank);
    PH_CHECK(hipMemcpyAsync(c.data(), e->cols.p, slots * sizeof(int), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipMemcpyAsync(v.data(), e->vals.p, slots * sizeof(double), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipMemcpyAsync(n.data(), e->cnt.p, e->rank * sizeof(int), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipMemcpyAsync(r.data(), e->rhs.p, e->rank * sizeof(double), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipStreamSynchronize(st));
    // fresh rows: the reference's 5 slots (-1 = empty) or more when a row
    // grew; rebuilt rows: exactly their entries
    int total = 0;
    for (int i = 0; i < e->rank; ++i) total += e->compact ? n[i] : std::max(5, n[i]);
    if (!rowPtr) {
      *nslots = total;
      return;
    }
    if (*nslots < total) fail(PHASE_HIP_EINVAL, "output arrays too small");
    int w = 0;
    rowPtr[0] = 0;
    for (int i = 0; i < e->rank; ++i) {
      const int width = e->compact ? n[i] : std::max(5, n[i]);
      for (int j = 0; j < width; ++j, ++w) {
        const bool used = j < n[i];
        colInd[w] = used ? c[(size_t)i * e->stride + j] : -1;
        vals[w] = used ? v[(size_t)i * e->stride + j] : 0.;
      }
      rowPtr[i + 1] = w;
      if (rhs) rhs[i] = r[i];
    }
    *nslots = total;
  });
}

int phase_hip_solver_set_equation(phase_hip_solver *s, const phase_hip_equation *e) {
  if (!s || !e) return PHASE_HIP_EINVAL;
  try {
    // scalar equations carry one row per cell: their centroids order the rows
    const bool scalar = e->ncomp == 1;
    solverSetFromRows(s, e->rank, e->stride, e->cols.p, e->vals.p, e->rhs.p, e->mesh->stream,
                      scalar ? e->mesh->dCC.p : nullptr, scalar ? e->mesh->iptr.data() : nullptr,
                      scalar ? e->mesh->inb.data() : nullptr);
    return PHASE_HIP_OK;
  } catch (const HipError &x) {
    s->err = x.msg;
    return x.code;
  } catch (const std::exception &x) {
    s->err = x.what();
    return PHASE_HIP_EINVAL;
  }
}

int phase_hip_field_src_div(phase_hip_mesh *m, const double *uFace, double *out) {
  return guardedM(m, [&] {
    if (!uFace || !out) fail(PHASE_HIP_EINVAL, "null field");
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kSrcDiv, dim3(g256(m->nCells)), dim3(256), 0, m->stream, meshDev(m), uFace, out);
    PH_CHECK(hipGetLastError());
  });
}

int phase_hip_field_src_src(phase_hip_mesh *m, int ncomp, const double *phiCell, double *out) {
  return guardedM(m, [&] {
    if (!phiCell || !out || (ncomp != 1 && ncomp != 2)) fail(PHASE_HIP_EINVAL, "bad field");
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kSrcSrc, dim3(g256(m->nCells)), dim3(256), 0, m->stream, meshDev(m), ncomp, phiCell, out);
    PH_CHECK(hipGetLastError());
  });
}

int phase_hip_field_grad(phase_hip_mesh *m, const double *phiCell, const double *phiFace, double *gradFace,
                         double *gradCell) {
  return guardedM(m, [&] {
    if (!phiCell || !phiFace || !gradFace || !gradCell) fail(PHASE_HIP_EINVAL, "null field");
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kGradFaces, dim3(g256(m->nFaces)), dim3(256), 0, m->stream, m->nFaces, m->dFL.p, m->dFR.p,
                       m->dCC.p, m->dFC.p, phiCell, phiFace, gradFace);
    hipLaunchKernelGGL(kGradCells, dim3(g256(m->nCells)), dim3(256), 0, m->stream, meshDev(m), gradFace, gradCell);
    PH_CHECK(hipGetLastError());
  });
}

int phase_hip_field_interpolate_faces(phase_hip_mesh *m, int ncomp, const double *phiCell, const int *bcType,
                                      double *phiFace) {
  return guardedM(m, [&] {
    if (!phiCell || !phiFace || (ncomp != 1 && ncomp != 2)) fail(PHASE_HIP_EINVAL, "bad field");
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kInterpFaces, dim3(g256(m->nFaces)), dim3(256), 0, m->stream, m->nFaces, ncomp, m->dFL.p,
                       m->dFR.p, m->dCC.p, m->dFC.p, m->dSf.p, bcType, phiCell, phiFace);
    PH_CHECK(hipGetLastError());
  });
}

int phase_hip_field_face_to_cell(phase_hip_mesh *m, const double *faceVec, const double *cellWeight,
                                 const double *faceWeight, double *cellVec) {
  return guardedM(m, [&] {
    if (!faceVec || !cellVec) fail(PHASE_HIP_EINVAL, "null field");
    if (!cellWeight != !faceWeight) fail(PHASE_HIP_EINVAL, "faceToCell takes both weights or neither");
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kFaceToCell, dim3(g256(m->nCells)), dim3(256), 0, m->stream, meshDev(m), faceVec, cellWeight,
                       faceWeight, cellVec);
    PH_CHECK(hipGetLastError());
  });
}

int phase_hip_cicsam_weights(phase_hip_mesh *m, const double *uFace, const double *gammaCell,
                             const double *gradGammaCell, double timeStep, double *beta) {
  return guardedM(m, [&] {
    if (!uFace || !gammaCell || !gradGammaCell || !beta) fail(PHASE_HIP_EINVAL, "null field");
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kCicsamWeights, dim3(g256(m->nFaces)), dim3(256), 0, m->stream, meshDev(m), m->nFaces,
                       m->dFL.p, m->dFR.p, uFace, gammaCell, gradGammaCell, timeStep, beta);
    PH_CHECK(hipGetLastError());
  });
}

int phase_hip_cicsam_momentum_flux(phase_hip_mesh *m, double rho1, double rho2, const double *uFace,
                                   const double *gammaCell, const double *gammaFace, const double *beta,
                                   double *rhoU) {
  return guardedM(m, [&] {
    if (!uFace || !gammaCell || !gammaFace || !beta || !rhoU) fail(PHASE_HIP_EINVAL, "null field");
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kCicsamMomentumFlux, dim3(g256(m->nFaces)), dim3(256), 0, m->stream, meshDev(m), m->nFaces,
                       m->dFL.p, m->dFR.p, rho1, rho2, uFace, gammaCell, gammaFace, beta, rhoU);
    PH_CHECK(hipGetLastError());
  });
}

int phase_hip_vof_properties(phase_hip_mesh *m, double rho1, double rho2, double mu1, double mu2,
                             const double *gammaCell, const double *gammaFace, double *rhoCell, double *rhoFace,
                             double *muCell, double *muFace) {
  return guardedM(m, [&] {
    if (!gammaCell || !gammaFace || !rhoCell || !rhoFace || !muCell || !muFace) fail(PHASE_HIP_EINVAL, "null field");
    if (!(mu1 > 0.) || !(mu2 > 0.)) fail(PHASE_HIP_EINVAL, "viscosities must be positive");
    PH_CHECK(hipSetDevice(m->device));
    hipLaunchKernelGGL(kVofProperties, dim3(g256(m->nCells)), dim3(256), 0, m->stream, m->nCells, rho1, rho2, mu1,
                       mu2, gammaCell, rhoCell, muCell);
    hipLaunchKernelGGL(kVofProperties, dim3(g256(m->nFaces)), dim3(256), 0, m->stream, m->nFaces, rho1, rho2, mu1,
                       mu2, gammaFace, rhoFace, muFace);
    PH_CHECK(hipGetLastError());
  });
}

int phase_hip_mesh_synchronize(phase_hip_mesh *m) {
  return guardedM(m, [&] { PH_CHECK(hipStreamSynchronize(m->stream)); });
}

}  // extern "C"
