// Outer-loop drivers on the device: the reference's fractional-step
// projection solver (FractionalStep, below) and a SIMPLE driver (Patankar &
// Spalding) for steady incompressible flow: the momentum and pressure-correction equations of the reference's
// README (README.md:27-37), assembled with the device fv:: operators and solved
// through the device solver plugin, with collocated Rhie-Chow face velocities.
// The reference snapshot ships no SIMPLE module (its solvers are projection
// methods, src/2D/Unstructured/Solvers/FractionalStep.cpp); this follows the
// README equations and the textbook algorithm, with the implicit
// under-relaxation that survives in the comment of
// ScalarFiniteVolumeEquation.cpp:45-55.
//
// One outer iteration, all fields in HBM:
//   p_f = setBoundaryFaces/interpolateFaces(p);  grad p (ScalarGradient)
//   uEqn = rho div(u_f, u) == laplacian(mu, u) - src(grad p);  relax(alpha_u)
//   d_P = V_P / a_P (relaxed diagonal);  solve uEqn -> u*
//   u_f = avg(u*) + avg(d) (avg(grad p) - grad_f p)       (Rhie-Chow)
//   pCorrEqn = laplacian(rho avg(d), p') == src::div(u_f), one cell pinned to 0
//   u -= d grad p';  u_f -= avg(d) grad_f p';  p += alpha_p p'
// Boundaries: velocity FIXED faces take their prescribed value; pressure and
// pressure correction are zero-gradient (NORMAL_GRADIENT) everywhere, so the
// pressure-correction matrix is singular and the `pin` cell's row is replaced
// by p'_pin = 0.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "celeste.h"
#include "equation.h"
#include "mesh.h"
#include "solver.h"
#include "solver_kernels.h"

using namespace phase;

struct phase_hip_simple {
  phase_hip_mesh *mesh = nullptr;
  int N = 0, F = 0, pin = 0;
  double rho = 1., mu = 1., alphaU = 0.7, alphaP = 0.3;
  phase_hip_equation *uEq = nullptr, *tmp = nullptr, *pEq = nullptr;
  phase_hip_solver *su = nullptr, *sp = nullptr;
  DBuf<double> uC, u0C, uB, uF, uWall, pC, pF, gpF, gpC, gsrc, d, df, rdf, m, pp, ppF, gppF, gppC, zeroF, maxBuf;
  DBuf<int> uBc, pBc;
  std::string err;
};

namespace {

inline int g256(long n) { return (int)std::max<long>(1, std::min<long>((n + 255) / 256, 8192)); }

// interleaved (x0, y0, x1, y1, ...) <-> block (x0..xN-1, y0..yN-1)
__global__ void kToBlocks(int n, const double *__restrict__ in, double *__restrict__ out) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    out[c] = in[2 * c];
    out[n + c] = in[2 * c + 1];
  }
}
__global__ void kFromBlocks(int n, const double *__restrict__ in, double *__restrict__ out) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    out[2 * c] = in[c];
    out[2 * c + 1] = in[n + c];
  }
}

// d_P = V_P / a_P from the diagonal of the x-momentum row of cell P
__global__ void kMomentumD(int n, int stride, const int *__restrict__ cols, const double *__restrict__ vals,
                           const int *__restrict__ cnt, const double *__restrict__ vol, double *d) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    double ap = 0.;
    for (int j = 0; j < cnt[c]; ++j)
      if (cols[(size_t)c * stride + j] == c) ap = vals[(size_t)c * stride + j];
    d[c] = vol[c] / ap;
  }
}

// Rhie-Chow face velocity; rdf = rho * avg(d) (pressure-correction
// coefficient), df = avg(d).  Face weights are Face::distanceWeight
// (Face.cpp:60-64), as in interpolateFaces.
__global__ void kRhieChow(int nFaces, const int *__restrict__ fl, const int *__restrict__ fr,
                          const double *__restrict__ cc, const double *__restrict__ fc,
                          const double *__restrict__ uC, const double *__restrict__ d,
                          const double *__restrict__ gpC, const double *__restrict__ gpF,
                          const double *__restrict__ uWall, double rho, double *uF, double *df, double *rdf) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nFaces; f += gridDim.x * blockDim.x) {
    const int l = fl[f], r = fr[f];
    if (r < 0) {
      uF[2 * f] = uWall[2 * f];
      uF[2 * f + 1] = uWall[2 * f + 1];
      df[f] = d[l];
      rdf[f] = rho * d[l];
      continue;
    }
    const double ax = fc[2 * f] - cc[2 * r], ay = fc[2 * f + 1] - cc[2 * r + 1];
    const double bx = fc[2 * f] - cc[2 * l], by = fc[2 * f + 1] - cc[2 * l + 1];
    const double l1 = sqrt(ax * ax + ay * ay), l2 = sqrt(bx * bx + by * by);
    const double g = l1 / (l1 + l2);
    const double dd = g * d[l] + (1. - g) * d[r];
    for (int q = 0; q < 2; ++q) {
      const double ub = g * uC[2 * l + q] + (1. - g) * uC[2 * r + q];
      const double gb = g * gpC[2 * l + q] + (1. - g) * gpC[2 * r + q];
      uF[2 * f + q] = ub + dd * (gb - gpF[2 * f + q]);
    }
    df[f] = dd;
    rdf[f] = rho * dd;
  }
}

// row `pin` of a scalar equation becomes x_pin = 0
// Row `pin` becomes a_pp p_pin = 0 with the row's own diagonal a_pp (1 when
// it has none): the same solution as a unit row.  Under ILU(0) the Krylov
// iterates are the same either way (the pin component of every residual stays
// 0); the multigrid's Galerkin sums add the row into its aggregate, where a
// row of the laplacian's sign and size belongs.
__global__ void kPinRow(int pin, int stride, int *cols, double *vals, int *cnt, double *rhs) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double d = 0.;
    for (int j = 0; j < cnt[pin]; ++j)
      if (cols[(size_t)pin * stride + j] == pin) d += vals[(size_t)pin * stride + j];
    cols[(size_t)pin * stride] = pin;
    vals[(size_t)pin * stride] = d != 0. ? d : 1.;
    for (int j = 1; j < stride; ++j) {
      cols[(size_t)pin * stride + j] = -1;
      vals[(size_t)pin * stride + j] = 0.;
    }
    cnt[pin] = 1;
    rhs[pin] = 0.;
  }
}

// u_P -= d_P grad p'_P ; p += alpha_p p'
__global__ void kCorrectCells(int n, const double *__restrict__ d, const double *__restrict__ gppC,
                              const double *__restrict__ pp, double alphaP, double *uC, double *pC) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    uC[2 * c] -= d[c] * gppC[2 * c];
    uC[2 * c + 1] -= d[c] * gppC[2 * c + 1];
    pC[c] += alphaP * pp[c];
  }
}

// interior faces: u_f -= avg(d) grad_f p'
__global__ void kCorrectFaces(int nFaces, const int *__restrict__ fr, const double *__restrict__ df,
                              const double *__restrict__ gppF, double *uF) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nFaces; f += gridDim.x * blockDim.x) {
    if (fr[f] < 0) continue;
    uF[2 * f] -= df[f] * gppF[2 * f];
    uF[2 * f + 1] -= df[f] * gppF[2 * f + 1];
  }
}

// u += s * g (both interleaved vectors of n entries)
__global__ void kAxpy(long n, double sc, const double *__restrict__ g, double *u) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) u[i] += sc * g[i];
}

// max |a_i| in two passes: block maxima into out[1 + b] (fixed order within a
// block), then one block takes the maximum of those into out[0]
constexpr int kMaxBlocks = 1024;
__global__ void kMaxAbs(int n, const double *__restrict__ a, double *out) {
  __shared__ double s[256];
  double m = 0.;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) m = fmax(m, fabs(a[i]));
  s[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[1 + blockIdx.x] = s[0];
}
__global__ void kMaxAbsFinish(int nb, double *out) {
  __shared__ double s[256];
  double m = 0.;
  for (int i = threadIdx.x; i < nb; i += 256) m = fmax(m, out[1 + i]);
  s[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}
inline int maxAbsBlocks(int n) { return std::max(1, std::min(kMaxBlocks, (n + 255) / 256)); }

void ck(int rc, const char *what, const char *detail) {
  if (rc != PHASE_HIP_OK) fail(rc, std::string(what) + ": " + (detail ? detail : ""));
}

// The value of the numPressureCorrections key (the name the reference's case
// files use, Examples/RisingBubble/case/case.info:16): 1 to 10 passes of
// momentum predictor, pressure solve and velocity correction per time step.
int parseCorrections(const char *value) {
  char *end = nullptr;
  const long n = std::strtol(value, &end, 10);
  if (end == value || *end != '\0' || n < 1 || n > 10)
    fail(PHASE_HIP_EINVAL, "numPressureCorrections takes an integer from 1 to 10");
  return (int)n;
}

template <class F>
int guardedS(phase_hip_simple *s, F &&f) {
  if (!s) return PHASE_HIP_EINVAL;
  try {
    f();
    return PHASE_HIP_OK;
  } catch (const HipError &x) {
    s->err = x.msg;
    return x.code;
  } catch (const std::exception &x) {
    s->err = x.what();
    return PHASE_HIP_EINVAL;
  }
}

double maxAbs(phase_hip_simple *s, const double *a, int n) {
  hipLaunchKernelGGL(kMaxAbs, dim3(maxAbsBlocks(n)), dim3(256), 0, s->mesh->stream, n, a, s->maxBuf.p);
  hipLaunchKernelGGL(kMaxAbsFinish, dim3(1), dim3(256), 0, s->mesh->stream, maxAbsBlocks(n), s->maxBuf.p);
  double v = 0.;
  PH_CHECK(hipMemcpyAsync(&v, s->maxBuf.p, sizeof(double), hipMemcpyDeviceToHost, s->mesh->stream));
  PH_CHECK(hipStreamSynchronize(s->mesh->stream));
  return v;
}

// p_f (zero gradient on the boundary) and grad p, faces and cells
void pressureGradient(phase_hip_simple *s) {
  phase_hip_mesh *m = s->mesh;
  ck(phase_hip_field_interpolate_faces(m, 1, s->pC.p, s->pBc.p, s->pF.p), "p faces", m->err.c_str());
  ck(phase_hip_field_grad(m, s->pC.p, s->pF.p, s->gpF.p, s->gpC.p), "grad p", m->err.c_str());
}

// uEqn = rho (ddt(u) +) rho div(u_f, u) == laplacian(mu, u) - src(grad p);
// dt <= 0: steady (no ddt); alphaU < 1: implicit relaxation.  Leaves d = V/a_P
// and the solved velocity in uC; returns the momentum iterations.
int momentum(phase_hip_simple *s, double dt, double alphaU) {
  phase_hip_mesh *m = s->mesh;
  hipStream_t st = m->stream;
  const int N = s->N;
  ck(phase_hip_field_src_src(m, 2, s->gpC.p, s->gsrc.p), "src(grad p)", m->err.c_str());
  ck(phase_hip_eq_div(s->uEq, s->uF.p, s->uF.p, 1.0, s->uBc.p, s->uWall.p, s->uC.p, s->uWall.p), "fv::div",
     s->uEq->err.c_str());
  if (dt > 0.) {  // fv::ddt(rho, u, dt) + fv::div(rho u, u), TimeDerivative.h
    ck(phase_hip_eq_ddt(s->tmp, dt, s->u0C.p), "fv::ddt", s->tmp->err.c_str());
    ck(phase_hip_eq_add(s->tmp, s->uEq, 1), "operator+", s->tmp->err.c_str());
    ck(phase_hip_eq_copy(s->uEq, s->tmp), "copy", s->uEq->err.c_str());
  }
  if (s->rho != 1.) ck(phase_hip_eq_scale(s->uEq, s->rho), "scale", s->uEq->err.c_str());
  ck(phase_hip_eq_laplacian(s->tmp, 1, s->mu, nullptr, nullptr, 1.0, s->uBc.p, s->uWall.p, nullptr, nullptr),
     "fv::laplacian", s->tmp->err.c_str());
  ck(phase_hip_eq_sub_vector(s->tmp, s->gsrc.p), "operator-", s->tmp->err.c_str());
  ck(phase_hip_eq_add(s->uEq, s->tmp, -1), "operator==", s->uEq->err.c_str());
  hipLaunchKernelGGL(kToBlocks, dim3(g256(N)), dim3(256), 0, st, N, s->uC.p, s->uB.p);
  if (alphaU < 1.) ck(phase_hip_eq_relax(s->uEq, alphaU, s->uB.p), "relax", s->uEq->err.c_str());
  hipLaunchKernelGGL(kMomentumD, dim3(g256(N)), dim3(256), 0, st, N, s->uEq->stride, s->uEq->cols.p,
                     s->uEq->vals.p, s->uEq->cnt.p, m->dVol.p, s->d.p);
  // momentum solve from the current velocity; like the reference's solve(),
  // an unconverged inner solve returns its best iterate and the loop goes on
  ck(phase_hip_solver_set_equation(s->su, s->uEq), "uEqn set", s->su->err.c_str());
  ck(phase_hip_solver_set_guess_device(s->su, s->uB.p), "uEqn guess", s->su->err.c_str());
  double err = 0.;
  const int rc = phase_hip_solver_solve(s->su, &err);
  if (rc != PHASE_HIP_ENOCONV) ck(rc, "uEqn solve", s->su->err.c_str());
  ck(phase_hip_solver_get_x_device(s->su, s->uB.p), "uEqn x", s->su->err.c_str());
  hipLaunchKernelGGL(kFromBlocks, dim3(g256(N)), dim3(256), 0, st, N, s->uB.p, s->uC.p);
  return phase_hip_solver_n_iters(s->su);
}

// Rhie-Chow face velocities from uC and grad p, pCorrEqn = laplacian(rho d_f,
// p') == src::div(u_f) with p'_pin = 0, then u -= d grad p', u_f -= d_f
// grad_f p', p += alphaP p'.  Returns the pressure iterations.
int pressureCorrection(phase_hip_simple *s, double alphaP, double *mMax, double *ppMax) {
  phase_hip_mesh *m = s->mesh;
  hipStream_t st = m->stream;
  const int N = s->N, F = s->F;
  hipLaunchKernelGGL(kRhieChow, dim3(g256(F)), dim3(256), 0, st, F, m->dFL.p, m->dFR.p, m->dCC.p, m->dFC.p,
                     s->uC.p, s->d.p, s->gpC.p, s->gpF.p, s->uWall.p, s->rho, s->uF.p, s->df.p, s->rdf.p);
  ck(phase_hip_field_src_div(m, s->uF.p, s->m.p), "src::div", m->err.c_str());
  if (mMax) *mMax = maxAbs(s, s->m.p, N);
  ck(phase_hip_eq_laplacian(s->pEq, 3, 0., s->rdf.p, nullptr, 1.0, s->pBc.p, s->zeroF.p, nullptr, nullptr),
     "fv::laplacian", s->pEq->err.c_str());
  ck(phase_hip_eq_sub_vector(s->pEq, s->m.p), "operator==", s->pEq->err.c_str());
  hipLaunchKernelGGL(kPinRow, dim3(1), dim3(64), 0, st, s->pin, s->pEq->stride, s->pEq->cols.p, s->pEq->vals.p,
                     s->pEq->cnt.p, s->pEq->rhs.p);
  ck(phase_hip_solver_set_equation(s->sp, s->pEq), "pCorrEqn set", s->sp->err.c_str());
  ck(phase_hip_solver_zero_guess(s->sp), "pCorrEqn guess", s->sp->err.c_str());
  double err = 0.;
  const int rc = phase_hip_solver_solve(s->sp, &err);
  if (rc != PHASE_HIP_ENOCONV) ck(rc, "pCorrEqn solve", s->sp->err.c_str());
  ck(phase_hip_solver_get_x_device(s->sp, s->pp.p), "pCorrEqn x", s->sp->err.c_str());
  ck(phase_hip_field_interpolate_faces(m, 1, s->pp.p, s->pBc.p, s->ppF.p), "p' faces", m->err.c_str());
  ck(phase_hip_field_grad(m, s->pp.p, s->ppF.p, s->gppF.p, s->gppC.p), "grad p'", m->err.c_str());
  hipLaunchKernelGGL(kCorrectCells, dim3(g256(N)), dim3(256), 0, st, N, s->d.p, s->gppC.p, s->pp.p, alphaP,
                     s->uC.p, s->pC.p);
  hipLaunchKernelGGL(kCorrectFaces, dim3(g256(F)), dim3(256), 0, st, F, m->dFR.p, s->df.p, s->gppF.p, s->uF.p);
  PH_CHECK(hipGetLastError());
  if (ppMax) *ppMax = maxAbs(s, s->pp.p, N);
  return phase_hip_solver_n_iters(s->sp);
}

// One SIMPLE outer iteration (steady, relaxed).
void outerIteration(phase_hip_simple *s, double *hist) {
  pressureGradient(s);
  const int uIt = momentum(s, 0., s->alphaU);
  double mMax = 0., ppMax = 0.;
  const int pIt = pressureCorrection(s, s->alphaP, &mMax, &ppMax);
  if (hist) {
    hist[0] = uIt;
    hist[1] = pIt;
    hist[2] = mMax;
    hist[3] = ppMax;
  }
}

// One PISO time step (Issa 1986): momentum predictor with fv::ddt and no
// relaxation, then nCorr pressure corrections, each from the latest velocity
// and pressure gradient.
void pisoStep(phase_hip_simple *s, double dt, int nCorr, double *hist) {
  hipStream_t st = s->mesh->stream;
  PH_CHECK(hipMemcpyAsync(s->u0C.p, s->uC.p, 2 * (size_t)s->N * sizeof(double), hipMemcpyDeviceToDevice, st));
  pressureGradient(s);
  const int uIt = momentum(s, dt, 1.0);
  int pIt = 0;
  double mMax = 0., ppMax = 0.;
  for (int k = 0; k < nCorr; ++k) {
    if (k > 0) pressureGradient(s);
    pIt += pressureCorrection(s, 1.0, &mMax, &ppMax);
  }
  if (hist) {
    hist[0] = uIt;
    hist[1] = pIt;
    hist[2] = mMax;
    hist[3] = ppMax;
  }
}

}  // namespace

// FractionalStep::solve (src/2D/Unstructured/Solvers/FractionalStep.cpp:36-117)
// on the device, one time step:
//   u0 = u (cells and faces)                                   savePreviousTimeStep
//   uEqn = ddt(u, dt) + div(u, u, 0) == laplacian(mu/rho, u, 0.5) - src(grad p)
//   u = solve(uEqn);  u += dt grad p;  u_f = interpolateFaces(u)
//   pEqn = laplacian(dt, p) == src::div(u_f);  p = solve(pEqn)  (pin cell fixed at 0)
//   p_f = setBoundaryFaces(p);  grad p (faces, FACE_TO_CELL)
//   u -= dt grad p (cells);  u_f -= dt grad_f p (every face)
//
// With an immersed boundary (phase_hip_fstep_immersed_boundary) the step is
// FractionalStepGCIB::solve (src/2D/Unstructured/Solvers/FractionalStepGCIB.cpp:14-56):
// the fv:: terms keep the fluid cells' rows only, velocityBcs(u) and bcs(p)
// add the ghost and solid rows, the momentum source is src(grad p / rho),
// u += dt / rho grad p and u -= dt grad p touch the fluid cells only, and
// src::div and the divergence check cover the fluid cells.
//
// Setup key numPressureCorrections = n (default 1, the reference's single
// pass) repeats everything after savePreviousTimeStep n times per step: each
// later pass assembles the momentum equation with the same old-time fields
// but the latest face velocities and pressure gradient, then projects again
// (the fractional-step form of PISO's repeated correctors).  The reference's
// case files carry the key; its fractional-step solvers make one pass.
struct phase_hip_fstep {
  phase_hip_mesh *mesh = nullptr;
  int N = 0, F = 0, pin = 0;
  double rho = 1., mu = 1.;
  phase_hip_equation *uEq = nullptr, *tmp = nullptr, *pEq = nullptr;
  phase_hip_solver *su = nullptr, *sp = nullptr;
  DBuf<double> uC, u0C, uB, uF, u0F, pC, pB, pF, gpF, gpC, gsrc, m, maxBuf;
  DBuf<int> uBc, pBc;
  // immersed boundary: cell types (0 fluid), its momentum and pressure rows
  bool immersed = false;
  DBuf<int> ibTypes;
  DBuf<double> gq;
  phase_hip_equation *ibU = nullptr, *ibP = nullptr;
  double ibPScale = 1.;  // factor the ghost pressure rows carry (see timeStep)
  double ibH = 1.;       // the ghost cells' mean size sqrt(volume)
  int ibPBc = PHASE_HIP_BC_NORMAL_GRADIENT;
  int nCorr = 1;  // numPressureCorrections
  std::string err;
};

namespace {

template <class F>
int guardedF(phase_hip_fstep *s, F &&f) {
  if (!s) return PHASE_HIP_EINVAL;
  try {
    f();
    return PHASE_HIP_OK;
  } catch (const HipError &x) {
    s->err = x.msg;
    return x.code;
  } catch (const std::exception &x) {
    s->err = x.what();
    return PHASE_HIP_EINVAL;
  }
}

// u_c += sc g_c for the fluid cells (type 0); u, g interleaved
__global__ void kFluidAxpy(int n, double sc, const int *__restrict__ type, const double *__restrict__ g, double *u) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
    if (type[c] == 0) {
      u[2 * c] += sc * g[2 * c];
      u[2 * c + 1] += sc * g[2 * c + 1];
    }
}

// v = 0 on the rows of the non-fluid cells (ncomp blocks of n rows)
__global__ void kZeroNonFluid(int n, int ncomp, const int *__restrict__ type, double *v) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
    if (type[c] != 0)
      for (int q = 0; q < ncomp; ++q) v[(long)q * n + c] = 0.;
}

// out = a / d
__global__ void kDivide(long n, double d, const double *__restrict__ a, double *__restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = a[i] / d;
}

double maxAbsF(phase_hip_fstep *s, const double *a, int n) {
  hipLaunchKernelGGL(kMaxAbs, dim3(maxAbsBlocks(n)), dim3(256), 0, s->mesh->stream, n, a, s->maxBuf.p);
  hipLaunchKernelGGL(kMaxAbsFinish, dim3(1), dim3(256), 0, s->mesh->stream, maxAbsBlocks(n), s->maxBuf.p);
  double v = 0.;
  PH_CHECK(hipMemcpyAsync(&v, s->maxBuf.p, sizeof(double), hipMemcpyDeviceToHost, s->mesh->stream));
  PH_CHECK(hipStreamSynchronize(s->mesh->stream));
  return v;
}

void timeStep(phase_hip_fstep *s, double dt, double *hist) {
  phase_hip_mesh *m = s->mesh;
  hipStream_t st = m->stream;
  const int N = s->N, F = s->F;
  // savePreviousTimeStep
  PH_CHECK(hipMemcpyAsync(s->u0C.p, s->uC.p, 2 * (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, st));
  PH_CHECK(hipMemcpyAsync(s->u0F.p, s->uF.p, 2 * (size_t)F * sizeof(double), hipMemcpyDeviceToDevice, st));
  const bool ib = s->immersed;
  int uIters = 0, pIters = 0;
  for (int corr = 0; corr < s->nCorr; ++corr) {
    // solveUEqn (FractionalStep.cpp:79-92)
    ck(phase_hip_eq_ddt(s->uEq, dt, s->u0C.p), "fv::ddt", s->uEq->err.c_str());
    ck(phase_hip_eq_div(s->tmp, s->uF.p, s->u0F.p, 0.0, s->uBc.p, s->uF.p, s->u0C.p, s->u0F.p), "fv::div",
       s->tmp->err.c_str());
    ck(phase_hip_eq_add(s->uEq, s->tmp, 1), "operator+", s->uEq->err.c_str());
    if (ib) {  // + ib_.velocityBcs(u_), the fv:: rows on the fluid cells
      ck(phase_hip_eq_remove_cells(s->uEq, s->ibTypes.p), "remove", s->uEq->err.c_str());
      ck(phase_hip_eq_add(s->uEq, s->ibU, 1), "operator+", s->uEq->err.c_str());
    }
    ck(phase_hip_eq_laplacian(s->tmp, 0, s->mu / s->rho, nullptr, nullptr, 0.5, s->uBc.p, s->uF.p, s->u0C.p, s->u0F.p),
       "fv::laplacian", s->tmp->err.c_str());
    if (ib) {  // - src::src(gradP_ / rho_) over the fluid cells
      ck(phase_hip_eq_remove_cells(s->tmp, s->ibTypes.p), "remove", s->tmp->err.c_str());
      hipLaunchKernelGGL(kDivide, dim3(g256(2L * N)), dim3(256), 0, st, 2L * N, s->rho, s->gpC.p, s->gq.p);
      ck(phase_hip_field_src_src(m, 2, s->gq.p, s->gsrc.p), "src::src", m->err.c_str());
      hipLaunchKernelGGL(kZeroNonFluid, dim3(g256(N)), dim3(256), 0, st, N, 2, s->ibTypes.p, s->gsrc.p);
    } else {
      ck(phase_hip_field_src_src(m, 2, s->gpC.p, s->gsrc.p), "src::src", m->err.c_str());
    }
    ck(phase_hip_eq_sub_vector(s->tmp, s->gsrc.p), "operator-", s->tmp->err.c_str());
    ck(phase_hip_eq_add(s->uEq, s->tmp, -1), "operator==", s->uEq->err.c_str());
    hipLaunchKernelGGL(kToBlocks, dim3(g256(N)), dim3(256), 0, st, N, s->uC.p, s->uB.p);
    ck(phase_hip_solver_set_equation(s->su, s->uEq), "uEqn set", s->su->err.c_str());
    ck(phase_hip_solver_set_guess_device(s->su, s->uB.p), "uEqn guess", s->su->err.c_str());
    double err = 0.;
    int rc = phase_hip_solver_solve(s->su, &err);
    if (rc != PHASE_HIP_ENOCONV) ck(rc, "uEqn solve", s->su->err.c_str());
    ck(phase_hip_solver_get_x_device(s->su, s->uB.p), "uEqn x", s->su->err.c_str());
    hipLaunchKernelGGL(kFromBlocks, dim3(g256(N)), dim3(256), 0, st, N, s->uB.p, s->uC.p);
    if (ib)
      hipLaunchKernelGGL(kFluidAxpy, dim3(g256(N)), dim3(256), 0, st, N, dt / s->rho, s->ibTypes.p, s->gpC.p, s->uC.p);
    else
      hipLaunchKernelGGL(kAxpy, dim3(g256(2L * N)), dim3(256), 0, st, 2L * N, dt, s->gpC.p, s->uC.p);
    ck(phase_hip_field_interpolate_faces(m, 2, s->uC.p, s->uBc.p, s->uF.p), "interpolateFaces", m->err.c_str());
    // solvePEqn (FractionalStep.cpp:94-106; FractionalStepGCIB.cpp:44-56)
    ck(phase_hip_eq_laplacian(s->pEq, 1, ib ? dt / s->rho : dt, nullptr, nullptr, 1.0, s->pBc.p, s->pF.p, nullptr,
                              nullptr),
       "fv::laplacian", s->pEq->err.c_str());
    if (ib) {  // + ib_.bcs(p_)
      // The ghost and solid rows are constraints independent of dt / rho, and
      // a NORMAL_GRADIENT ghost row's coefficients are of order 1 / h, while a
      // fluid row's diagonal is about -4 dt / rho (cells of aspect ratio near
      // one) with positive neighbours.  Scaled to that size and sign (-4 dt h /
      // rho for NORMAL_GRADIENT rows, -4 dt / rho for FIXED ones) they have
      // the same solution, and the Krylov solve and the multigrid coarsening
      // converge several times faster: with BiCGStab and multigrid at 1M cells
      // 627 iterations unscaled, 166 at +4 dt / rho, 71 scaled so; at 4M cells
      // 5,156, 1,002 and 164.
      const double want = -4.0 * dt / s->rho * (s->ibPBc == PHASE_HIP_BC_NORMAL_GRADIENT ? s->ibH : 1.0);
      if (want != s->ibPScale) {
        ck(phase_hip_eq_scale(s->ibP, want / s->ibPScale), "operator*", s->ibP->err.c_str());
        s->ibPScale = want;
      }
      ck(phase_hip_eq_remove_cells(s->pEq, s->ibTypes.p), "remove", s->pEq->err.c_str());
      ck(phase_hip_eq_add(s->pEq, s->ibP, 1), "operator+", s->pEq->err.c_str());
    }
    ck(phase_hip_field_src_div(m, s->uF.p, s->m.p), "src::div", m->err.c_str());
    if (ib) hipLaunchKernelGGL(kZeroNonFluid, dim3(g256(N)), dim3(256), 0, st, N, 1, s->ibTypes.p, s->m.p);
    ck(phase_hip_eq_sub_vector(s->pEq, s->m.p), "operator==", s->pEq->err.c_str());
    if (s->pin >= 0)  // all-Neumann pressure: one cell fixed at 0
      hipLaunchKernelGGL(kPinRow, dim3(1), dim3(64), 0, st, s->pin, s->pEq->stride, s->pEq->cols.p, s->pEq->vals.p,
                         s->pEq->cnt.p, s->pEq->rhs.p);
    ck(phase_hip_solver_set_equation(s->sp, s->pEq), "pEqn set", s->sp->err.c_str());
    ck(phase_hip_solver_set_guess_device(s->sp, s->pC.p), "pEqn guess", s->sp->err.c_str());
    rc = phase_hip_solver_solve(s->sp, &err);
    if (rc != PHASE_HIP_ENOCONV) ck(rc, "pEqn solve", s->sp->err.c_str());
    ck(phase_hip_solver_get_x_device(s->sp, s->pC.p), "pEqn x", s->sp->err.c_str());
    ck(phase_hip_field_interpolate_faces(m, 1, s->pC.p, s->pBc.p, s->pF.p), "setBoundaryFaces", m->err.c_str());
    ck(phase_hip_field_grad(m, s->pC.p, s->pF.p, s->gpF.p, s->gpC.p), "grad p", m->err.c_str());
    // correctVelocity (FractionalStep.cpp:108-117)
    if (ib)
      hipLaunchKernelGGL(kFluidAxpy, dim3(g256(N)), dim3(256), 0, st, N, -dt, s->ibTypes.p, s->gpC.p, s->uC.p);
    else
      hipLaunchKernelGGL(kAxpy, dim3(g256(2L * N)), dim3(256), 0, st, 2L * N, -dt, s->gpC.p, s->uC.p);
    hipLaunchKernelGGL(kAxpy, dim3(g256(2L * F)), dim3(256), 0, st, 2L * F, -dt, s->gpF.p, s->uF.p);
    PH_CHECK(hipGetLastError());
    uIters += phase_hip_solver_n_iters(s->su);
    pIters += phase_hip_solver_n_iters(s->sp);
  }
  if (hist) {
    // FractionalStep::maxDivergenceError (FractionalStep.cpp:119-135) on the corrected faces
    ck(phase_hip_field_src_div(m, s->uF.p, s->m.p), "src::div", m->err.c_str());
    if (ib) hipLaunchKernelGGL(kZeroNonFluid, dim3(g256(N)), dim3(256), 0, st, N, 1, s->ibTypes.p, s->m.p);
    hist[0] = uIters;
    hist[1] = pIters;
    hist[2] = maxAbsF(s, s->m.p, N);
  }
}

}  // namespace

extern "C" {

int phase_hip_fstep_destroy(phase_hip_fstep *s) {
  if (!s) return PHASE_HIP_OK;
  phase_hip_equation_destroy(s->ibU);
  phase_hip_equation_destroy(s->ibP);
  phase_hip_equation_destroy(s->uEq);
  phase_hip_equation_destroy(s->tmp);
  phase_hip_equation_destroy(s->pEq);
  phase_hip_solver_destroy(s->su);
  phase_hip_solver_destroy(s->sp);
  delete s;
  return PHASE_HIP_OK;
}

const char *phase_hip_fstep_last_error(const phase_hip_fstep *s) { return s ? s->err.c_str() : "null driver"; }

int phase_hip_fstep_create(phase_hip_fstep **out, phase_hip_mesh *m, double rho, double mu, const int *uBcType,
                           const double *uFace0, const int *pBcType, const double *pFace0, int pinCell) {
  if (!out || !m || !uBcType || !uFace0) return PHASE_HIP_EINVAL;
  *out = nullptr;
  phase_hip_fstep *s = new phase_hip_fstep();
  s->mesh = m;
  s->N = m->nCells;
  s->F = m->nFaces;
  s->rho = rho;
  s->mu = mu;
  s->pin = pinCell;
  const int rc = guardedF(s, [&] {
    if (!(rho > 0.) || !(mu > 0.)) fail(PHASE_HIP_EINVAL, "rho and mu must be positive");
    if (pinCell < -1 || pinCell >= m->nCells) fail(PHASE_HIP_EINVAL, "pin cell out of range (-1: none)");
    PH_CHECK(hipSetDevice(m->device));
    hipStream_t st = m->stream;
    const int N = s->N, F = s->F;
    ck(phase_hip_equation_create(&s->uEq, m, 2), "uEqn", "");
    ck(phase_hip_equation_create(&s->tmp, m, 2), "uEqn", "");
    ck(phase_hip_equation_create(&s->pEq, m, 1), "pEqn", "");
    ck(phase_hip_solver_create(&s->su, m->device), "solver", "");
    ck(phase_hip_solver_create(&s->sp, m->device), "solver", "");
    for (phase_hip_solver *v : {s->su, s->sp}) {
      ck(phase_hip_solver_setup(v, "tolerance", "1e-8"), "setup", v->err.c_str());
      ck(phase_hip_solver_setup(v, "maxIters", "5000"), "setup", v->err.c_str());
    }
    for (DBuf<double> *b : {&s->uC, &s->u0C, &s->uB, &s->gpC, &s->gsrc}) b->alloc(2 * (size_t)N);
    for (DBuf<double> *b : {&s->pC, &s->m}) b->alloc(N);
    for (DBuf<double> *b : {&s->uF, &s->u0F, &s->gpF}) b->alloc(2 * (size_t)F);
    s->pF.alloc(F);
    s->maxBuf.alloc(1 + kMaxBlocks);
    for (DBuf<double> *b : {&s->uC, &s->pC, &s->gpC, &s->gpF}) b->zero(st);
    std::vector<int> ubc(uBcType, uBcType + F), pbc(F, PHASE_HIP_BC_NORMAL_GRADIENT);
    if (pBcType) pbc.assign(pBcType, pBcType + F);
    s->uBc.upload(ubc, st);
    s->pBc.upload(pbc, st);
    // u_f at rest: the boundary values (FIXED faces keep them; interpolateFaces writes the rest)
    PH_CHECK(hipMemcpyAsync(s->uF.p, uFace0, 2 * (size_t)F * sizeof(double), hipMemcpyHostToDevice, st));
    if (pFace0) PH_CHECK(hipMemcpyAsync(s->pF.p, pFace0, (size_t)F * sizeof(double), hipMemcpyHostToDevice, st));
    else s->pF.zero(st);
    // FractionalStep::initialize (FractionalStep.cpp:24-27)
    ck(phase_hip_field_interpolate_faces(m, 2, s->uC.p, s->uBc.p, s->uF.p), "interpolateFaces", m->err.c_str());
    PH_CHECK(hipStreamSynchronize(st));
  });
  if (rc != PHASE_HIP_OK) {
    phase_hip_fstep_destroy(s);
    return rc;
  }
  *out = s;
  return PHASE_HIP_OK;
}

int phase_hip_fstep_setup(phase_hip_fstep *s, const char *key, const char *value) {
  return guardedF(s, [&] {
    if (!key || !value) fail(PHASE_HIP_EINVAL, "null key/value");
    std::string k(key);
    if (k == "numPressureCorrections") {
      s->nCorr = parseCorrections(value);
      return;
    }
    phase_hip_solver *v = nullptr;
    if (k.rfind("u.", 0) == 0) v = s->su;
    else if (k.rfind("p.", 0) == 0) v = s->sp;
    else fail(PHASE_HIP_EINVAL, "keys are numPressureCorrections, u.<solver key> or p.<solver key>");
    ck(phase_hip_solver_setup(v, k.c_str() + 2, value), "setup", v->err.c_str());
  });
}

}  // extern "C"

namespace {
// FractionalStepGCIB from the next step on: the velocity rows come from
// `velocityRows` (translating or moving bodies), the pressure rows from bcs(p).
template <class VelocityRows>
int attachImmersedBoundary(phase_hip_fstep *s, const phase_hip_gcib *ib, VelocityRows &&velocityRows, int pBcType,
                           double pRefValue) {
  return guardedF(s, [&] {
    if (!ib) fail(PHASE_HIP_EINVAL, "null immersed boundary");
    phase_hip_mesh *m = s->mesh;
    PH_CHECK(hipSetDevice(m->device));
    if (!s->ibU) ck(phase_hip_equation_create(&s->ibU, m, 2), "velocityBcs", "");
    if (!s->ibP) ck(phase_hip_equation_create(&s->ibP, m, 1), "bcs", "");
    ck(velocityRows(s->ibU), "velocityBcs", s->ibU->err.c_str());
    ck(phase_hip_eq_gcib_bcs(s->ibP, ib, pBcType, pRefValue), "bcs", s->ibP->err.c_str());
    s->ibPScale = 1.;
    s->ibPBc = pBcType;
    s->ibTypes.alloc(s->N);
    ck(phase_hip_gcib_cell_types(ib, s->ibTypes.p, nullptr, 1), "cell types", "");
    {
      std::vector<int> types(s->N);
      std::vector<double> vol(s->N);
      ck(phase_hip_gcib_cell_types(ib, types.data(), nullptr, 0), "cell types", "");
      ck(phase_hip_mesh_geometry(m, vol.data(), nullptr, nullptr, nullptr), "geometry", m->err.c_str());
      double hs = 0.;
      long ng = 0;
      for (int c = 0; c < s->N; ++c)
        if (types[c] == 1) {
          hs += std::sqrt(vol[c]);
          ++ng;
        }
      s->ibH = ng ? hs / ng : 1.;
    }
    s->gq.alloc(2 * (size_t)s->N);
    PH_CHECK(hipStreamSynchronize(m->stream));
    s->immersed = true;
  });
}
}  // namespace

extern "C" {

int phase_hip_fstep_immersed_boundary(phase_hip_fstep *s, const phase_hip_gcib *ib, const double *objectVelocity,
                                      int pBcType, double pRefValue) {
  return attachImmersedBoundary(
      s, ib, [&](phase_hip_equation *e) { return phase_hip_eq_gcib_velocity_bcs(e, ib, objectVelocity); }, pBcType,
      pRefValue);
}

int phase_hip_fstep_immersed_boundary_motion(phase_hip_fstep *s, const phase_hip_gcib *ib, const double *objectMotion,
                                             int pBcType, double pRefValue) {
  return attachImmersedBoundary(
      s, ib, [&](phase_hip_equation *e) { return phase_hip_eq_gcib_motion_bcs(e, ib, objectMotion); }, pBcType,
      pRefValue);
}

int phase_hip_fstep_advance(phase_hip_fstep *s, double dt, int nSteps, double *history) {
  return guardedF(s, [&] {
    if (nSteps < 0 || !(dt > 0.)) fail(PHASE_HIP_EINVAL, "need dt > 0 and nSteps >= 0");
    PH_CHECK(hipSetDevice(s->mesh->device));
    for (int it = 0; it < nSteps; ++it) timeStep(s, dt, history ? history + 3 * (size_t)it : nullptr);
  });
}

int phase_hip_fstep_fields(phase_hip_fstep *s, double *uCell, double *pCell, double *uFace) {
  return guardedF(s, [&] {
    hipStream_t st = s->mesh->stream;
    if (uCell) PH_CHECK(hipMemcpyAsync(uCell, s->uC.p, 2 * (size_t)s->N * sizeof(double), hipMemcpyDeviceToHost, st));
    if (pCell) PH_CHECK(hipMemcpyAsync(pCell, s->pC.p, (size_t)s->N * sizeof(double), hipMemcpyDeviceToHost, st));
    if (uFace) PH_CHECK(hipMemcpyAsync(uFace, s->uF.p, 2 * (size_t)s->F * sizeof(double), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipStreamSynchronize(st));
  });
}

const char *phase_hip_simple_last_error(const phase_hip_simple *s) { return s ? s->err.c_str() : "null driver"; }

int phase_hip_simple_destroy(phase_hip_simple *s) {
  if (!s) return PHASE_HIP_OK;
  phase_hip_equation_destroy(s->uEq);
  phase_hip_equation_destroy(s->tmp);
  phase_hip_equation_destroy(s->pEq);
  phase_hip_solver_destroy(s->su);
  phase_hip_solver_destroy(s->sp);
  delete s;
  return PHASE_HIP_OK;
}

int phase_hip_simple_create(phase_hip_simple **out, phase_hip_mesh *m, double rho, double mu, double alphaU,
                            double alphaP, const int *uBcType, const double *uBoundary, int pinCell) {
  if (!out || !m || !uBcType || !uBoundary) return PHASE_HIP_EINVAL;
  *out = nullptr;
  phase_hip_simple *s = new phase_hip_simple();
  s->mesh = m;
  s->N = m->nCells;
  s->F = m->nFaces;
  s->rho = rho;
  s->mu = mu;
  s->alphaU = alphaU;
  s->alphaP = alphaP;
  s->pin = pinCell;
  const int rc = guardedS(s, [&] {
    if (!(rho > 0.) || !(mu > 0.) || !(alphaU > 0. && alphaU <= 1.) || !(alphaP > 0. && alphaP <= 1.))
      fail(PHASE_HIP_EINVAL, "SIMPLE needs rho, mu > 0 and relaxation factors in (0, 1]");
    if (pinCell < 0 || pinCell >= m->nCells) fail(PHASE_HIP_EINVAL, "pin cell out of range");
    for (int f = 0; f < m->nFaces; ++f)
      if (m->fR[f] < 0 && uBcType[f] != PHASE_HIP_BC_FIXED)
        fail(PHASE_HIP_EBOUNDARY, "this SIMPLE driver takes FIXED velocity on every boundary face");
    PH_CHECK(hipSetDevice(m->device));
    hipStream_t st = m->stream;
    const int N = s->N, F = s->F;
    ck(phase_hip_equation_create(&s->uEq, m, 2), "uEqn", "");
    ck(phase_hip_equation_create(&s->tmp, m, 2), "uEqn", "");
    ck(phase_hip_equation_create(&s->pEq, m, 1), "pCorrEqn", "");
    ck(phase_hip_solver_create(&s->su, m->device), "solver", "");
    ck(phase_hip_solver_create(&s->sp, m->device), "solver", "");
    for (phase_hip_solver *v : {s->su, s->sp}) {
      ck(phase_hip_solver_setup(v, "tolerance", "1e-8"), "setup", v->err.c_str());
      ck(phase_hip_solver_setup(v, "maxIters", "5000"), "setup", v->err.c_str());
    }
    s->uC.alloc(2 * (size_t)N);
    s->u0C.alloc(2 * (size_t)N);
    s->uB.alloc(2 * (size_t)N);
    s->gpC.alloc(2 * (size_t)N);
    s->gsrc.alloc(2 * (size_t)N);
    s->gppC.alloc(2 * (size_t)N);
    s->pC.alloc(N);
    s->d.alloc(N);
    s->m.alloc(N);
    s->pp.alloc(N);
    s->uF.alloc(2 * (size_t)F);
    s->uWall.alloc(2 * (size_t)F);
    s->gpF.alloc(2 * (size_t)F);
    s->gppF.alloc(2 * (size_t)F);
    s->pF.alloc(F);
    s->ppF.alloc(F);
    s->df.alloc(F);
    s->rdf.alloc(F);
    s->zeroF.alloc(F);
    s->maxBuf.alloc(1 + kMaxBlocks);
    for (DBuf<double> *b : {&s->uC, &s->pC, &s->pF, &s->zeroF, &s->pp}) b->zero(st);
    std::vector<int> ubc(uBcType, uBcType + F), pbc(F, PHASE_HIP_BC_NORMAL_GRADIENT);
    s->uBc.upload(ubc, st);
    s->pBc.upload(pbc, st);
    std::vector<double> wall(uBoundary, uBoundary + 2 * (size_t)F);
    for (int f = 0; f < F; ++f)
      if (m->fR[f] >= 0) wall[2 * f] = wall[2 * f + 1] = 0.;
    PH_CHECK(hipMemcpyAsync(s->uWall.p, wall.data(), wall.size() * sizeof(double), hipMemcpyHostToDevice, st));
    // initial face velocities: the boundary values, zero inside (u = 0 at rest)
    PH_CHECK(hipMemcpyAsync(s->uF.p, s->uWall.p, 2 * (size_t)F * sizeof(double), hipMemcpyDeviceToDevice, st));
    PH_CHECK(hipStreamSynchronize(st));
  });
  if (rc != PHASE_HIP_OK) {
    phase_hip_simple_destroy(s);
    return rc;
  }
  *out = s;
  return PHASE_HIP_OK;
}

int phase_hip_simple_setup(phase_hip_simple *s, const char *key, const char *value) {
  return guardedS(s, [&] {
    if (!key || !value) fail(PHASE_HIP_EINVAL, "null key/value");
    std::string k(key);
    phase_hip_solver *v = nullptr;
    if (k.rfind("u.", 0) == 0) v = s->su;
    else if (k.rfind("p.", 0) == 0) v = s->sp;
    else fail(PHASE_HIP_EINVAL, "keys are u.<solver key> or p.<solver key>");
    ck(phase_hip_solver_setup(v, k.c_str() + 2, value), "setup", v->err.c_str());
  });
}

int phase_hip_simple_iterate(phase_hip_simple *s, int nIters, double *history) {
  return guardedS(s, [&] {
    if (nIters < 0) fail(PHASE_HIP_EINVAL, "negative iteration count");
    PH_CHECK(hipSetDevice(s->mesh->device));
    for (int it = 0; it < nIters; ++it) outerIteration(s, history ? history + 4 * (size_t)it : nullptr);
  });
}

int phase_hip_simple_piso_advance(phase_hip_simple *s, double dt, int nSteps, int nCorrectors, double *history) {
  return guardedS(s, [&] {
    if (nSteps < 0 || !(dt > 0.) || nCorrectors < 1) fail(PHASE_HIP_EINVAL, "need dt > 0, nSteps >= 0, nCorrectors >= 1");
    PH_CHECK(hipSetDevice(s->mesh->device));
    for (int it = 0; it < nSteps; ++it) pisoStep(s, dt, nCorrectors, history ? history + 4 * (size_t)it : nullptr);
  });
}

int phase_hip_simple_fields(phase_hip_simple *s, double *uCell, double *pCell, double *uFace) {
  return guardedS(s, [&] {
    hipStream_t st = s->mesh->stream;
    if (uCell) PH_CHECK(hipMemcpyAsync(uCell, s->uC.p, 2 * (size_t)s->N * sizeof(double), hipMemcpyDeviceToHost, st));
    if (pCell) PH_CHECK(hipMemcpyAsync(pCell, s->pC.p, (size_t)s->N * sizeof(double), hipMemcpyDeviceToHost, st));
    if (uFace) PH_CHECK(hipMemcpyAsync(uFace, s->uF.p, 2 * (size_t)s->F * sizeof(double), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipStreamSynchronize(st));
  });
}

}  // extern "C"

// FractionalStepMultiphase::solve (src/2D/Unstructured/Solvers/FractionalStepMultiphase.cpp:58-72)
// on the device, one time step (fst: the CELESTE surface-tension force when
// turned on, else 0):
//   solveGammaEqn (:76-100): beta = cicsam weights; ddt(gamma) + cicsam::div(u,
//     gamma, beta, 0.5) == 0; interpolateFaces(gamma); grad gamma
//   updateProperties (:175-218): rho0, mu0 = rho, mu; rho, mu from gamma;
//     grad rho (faces); sg_f = dot(g, -x_f) grad rho_f; sg = faceToCell(rho, rho);
//     fst_f = sigma kappa_f grad gamma_f; fst = faceToCell(rho, rho)
//   solveUEqn (:102-150): u1, u0 = u0, u; gradP = faceToCell(rho, rho0);
//     rho ddt(u) + rho dive(u, u, 0.5) == laplacian(mu, u, 0.5) + src(fst + sg - gradP);
//     u += dt / rho gradP; momentum-interpolated face velocities
//   solvePEqn (:152-163): laplacian(dt / rho, p) == src::div(u); boundary
//     faces; grad_f p; gradP = faceToCell(rho, rho)
//   correctVelocity (:165-173): u -= dt / rho gradP (cells and every face)
// numPressureCorrections = n repeats solveUEqn (after u1, u0 = u0, u),
// solvePEqn and correctVelocity n times per step, as in phase_hip_fstep.
struct phase_hip_multiphase {
  phase_hip_mesh *mesh = nullptr;
  int N = 0, F = 0, pin = -1;
  double rho1 = 1., rho2 = 1., mu1 = 1., mu2 = 1., gx = 0., gy = 0.;
  phase_hip_equation *gEq = nullptr, *uEq = nullptr, *tmp = nullptr, *gTmp = nullptr, *pEq = nullptr;
  phase_hip_solver *sg = nullptr, *su = nullptr, *sp = nullptr;
  DBuf<double> gC, gF, g0C, g0F, ggF, ggC, beta;                      // volume fraction
  DBuf<double> rhoC, rhoF, muC, muF, rho0F, mu0F, grF, grC, sgF, sgC;  // properties, gravity
  DBuf<double> uC, u0C, u1C, uF, u0F, u1F, uB, pC, pF, gpF, gpC, mob, gsrc, srcT, m, maxBuf;
  DBuf<double> fstF, fstC;  // surface tension (0 unless turned on)
  DBuf<int> uBc, pBc, gBc;
  phase::CelesteModel *st = nullptr;
  int nCorr = 1;  // numPressureCorrections
  std::string err;
};

namespace {

// sg_f = dot(g, -x_f) grad rho_f (FractionalStepMultiphase.cpp:190-192)
__global__ void kGravityFaces(int nFaces, double gx, double gy, const double *__restrict__ fc,
                              const double *__restrict__ grF, double *__restrict__ sgF) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nFaces; f += gridDim.x * blockDim.x) {
    const double h = gx * -fc[2 * f] + gy * -fc[2 * f + 1];
    sgF[2 * f] = h * grF[2 * f];
    sgF[2 * f + 1] = h * grF[2 * f + 1];
  }
}

// u_i += (dt / rho_c) g_i (add) or u_i -= (dt / rho_c) g_i, c = i / 2
__global__ void kDtOverRhoUpdate(long n2, double dt, int add, const double *__restrict__ rho,
                                 const double *__restrict__ g, double *u) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n2; i += (long)gridDim.x * blockDim.x) {
    const double v = (dt / rho[i >> 1]) * g[i];
    u[i] = add ? u[i] + v : u[i] - v;
  }
}

// out = dt / rho (the field timeStep / rho_ of solvePEqn)
__global__ void kScalarOver(int n, double a, const double *__restrict__ x, double *__restrict__ out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = a / x[i];
}

// out = gradP - (fst + sg): the negated cell source of the momentum equation
__global__ void kMomentumSource(long n, const double *__restrict__ gp, const double *__restrict__ fst,
                                const double *__restrict__ sg, double *__restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = gp[i] - (fst[i] + sg[i]);
}

// Face velocities after the momentum solve (FractionalStepMultiphase.cpp:118-147):
// interior faces volume-weighted (Face::volumeWeight = V_r / (V_r + V_l),
// Face.cpp:54-58) u - dt / rho (fst + sg) plus dt / rho_f (fst_f + sg_f); NORMAL_GRADIENT
// patches from the lCell; SYMMETRY patches the lCell's tangential component
// (Vector2D.cpp:30-36); FIXED (and OUTFLOW) faces unchanged.
__global__ void kMultiphaseFaceVelocity(int nFaces, double dt, const int *__restrict__ fl, const int *__restrict__ fr,
                                        const double *__restrict__ vol, const double *__restrict__ sf,
                                        const int *__restrict__ bc, const double *__restrict__ uC,
                                        const double *__restrict__ rhoC, const double *__restrict__ rhoF,
                                        const double *__restrict__ fstC, const double *__restrict__ fstF,
                                        const double *__restrict__ sgC, const double *__restrict__ sgF, double *uF) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nFaces; f += gridDim.x * blockDim.x) {
    const int a = fl[f], b = fr[f];
    const double kf = dt / rhoF[f];
    if (b >= 0) {
      const double v1 = vol[b], v2 = vol[a];
      const double g = v1 / (v1 + v2);
      const double ka = dt / rhoC[a], kb = dt / rhoC[b];
      for (int q = 0; q < 2; ++q)
        uF[2 * f + q] = g * (uC[2 * a + q] - ka * (fstC[2 * a + q] + sgC[2 * a + q])) +
                        (1. - g) * (uC[2 * b + q] - kb * (fstC[2 * b + q] + sgC[2 * b + q])) +
                        kf * (fstF[2 * f + q] + sgF[2 * f + q]);
    } else if (bc[f] == PHASE_HIP_BC_NORMAL_GRADIENT) {
      const double ka = dt / rhoC[a];
      for (int q = 0; q < 2; ++q)
        uF[2 * f + q] = uC[2 * a + q] - ka * (fstC[2 * a + q] + sgC[2 * a + q]) + kf * (fstF[2 * f + q] + sgF[2 * f + q]);
    } else if (bc[f] == PHASE_HIP_BC_SYMMETRY) {
      const double nx = sf[2 * f], ny = sf[2 * f + 1];
      const double t = (uC[2 * a] * nx + uC[2 * a + 1] * ny) / (nx * nx + ny * ny);
      uF[2 * f] = uC[2 * a] - t * nx;
      uF[2 * f + 1] = uC[2 * a + 1] - t * ny;
    }
  }
}

template <class Fn>
int guardedMP(phase_hip_multiphase *s, Fn &&f) {
  if (!s) return PHASE_HIP_EINVAL;
  try {
    f();
    return PHASE_HIP_OK;
  } catch (const HipError &x) {
    s->err = x.msg;
    return x.code;
  } catch (const std::exception &x) {
    s->err = x.what();
    return PHASE_HIP_EINVAL;
  }
}

double mpMaxAbs(phase_hip_multiphase *s, const double *a, int n) {
  hipLaunchKernelGGL(kMaxAbs, dim3(maxAbsBlocks(n)), dim3(256), 0, s->mesh->stream, n, a, s->maxBuf.p);
  hipLaunchKernelGGL(kMaxAbsFinish, dim3(1), dim3(256), 0, s->mesh->stream, maxAbsBlocks(n), s->maxBuf.p);
  double v = 0.;
  PH_CHECK(hipMemcpyAsync(&v, s->maxBuf.p, sizeof(double), hipMemcpyDeviceToHost, s->mesh->stream));
  PH_CHECK(hipStreamSynchronize(s->mesh->stream));
  return v;
}

void copyD(double *dst, const double *src, size_t n, hipStream_t st) {
  PH_CHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, st));
}

// solve e into x (device, original numbering), x also the initial guess;
// returns the iterations
int solveInto(phase_hip_solver *sv, phase_hip_equation *e, double *x, const char *what) {
  ck(phase_hip_solver_set_equation(sv, e), what, sv->err.c_str());
  ck(phase_hip_solver_set_guess_device(sv, x), what, sv->err.c_str());
  double err = 0.;
  const int rc = phase_hip_solver_solve(sv, &err);
  if (rc != PHASE_HIP_ENOCONV) ck(rc, what, sv->err.c_str());
  ck(phase_hip_solver_get_x_device(sv, x), what, sv->err.c_str());
  return phase_hip_solver_n_iters(sv);
}

// CELESTE force from the current gamma (Celeste::computeFaceInterfaceForces)
// and its cell values faceToCell(rho, rho) (FractionalStepMultiphase.cpp:211-214)
void mpSurfaceForce(phase_hip_multiphase *s) {
  phase_hip_mesh *m = s->mesh;
  phase::celesteFaceForces(m, *s->st, s->gC.p, s->ggF.p, s->fstF.p);
  ck(phase_hip_field_face_to_cell(m, s->fstF.p, s->rhoC.p, s->rhoF.p, s->fstC.p), "fst faceToCell", m->err.c_str());
}

// updateProperties (FractionalStepMultiphase.cpp:175-218)
void mpProperties(phase_hip_multiphase *s) {
  phase_hip_mesh *m = s->mesh;
  hipStream_t st = m->stream;
  copyD(s->rho0F.p, s->rhoF.p, s->F, st);
  copyD(s->mu0F.p, s->muF.p, s->F, st);
  ck(phase_hip_vof_properties(m, s->rho1, s->rho2, s->mu1, s->mu2, s->gC.p, s->gF.p, s->rhoC.p, s->rhoF.p, s->muC.p,
                              s->muF.p),
     "properties", m->err.c_str());
  ck(phase_hip_field_grad(m, s->rhoC.p, s->rhoF.p, s->grF.p, s->grC.p), "grad rho", m->err.c_str());
  hipLaunchKernelGGL(kGravityFaces, dim3(g256(s->F)), dim3(256), 0, st, s->F, s->gx, s->gy, m->dFC.p, s->grF.p,
                     s->sgF.p);
  ck(phase_hip_field_face_to_cell(m, s->sgF.p, s->rhoC.p, s->rhoF.p, s->sgC.p), "sg faceToCell", m->err.c_str());
  if (s->st) mpSurfaceForce(s);
}

void mpStep(phase_hip_multiphase *s, double dt, double *hist) {
  phase_hip_mesh *m = s->mesh;
  hipStream_t st = m->stream;
  const int N = s->N, F = s->F;
  // solveGammaEqn
  ck(phase_hip_cicsam_weights(m, s->uF.p, s->gC.p, s->ggC.p, dt, s->beta.p), "cicsam weights", m->err.c_str());
  copyD(s->g0C.p, s->gC.p, N, st);
  copyD(s->g0F.p, s->gF.p, F, st);
  ck(phase_hip_eq_ddt(s->gEq, dt, s->g0C.p), "fv::ddt", s->gEq->err.c_str());
  ck(phase_hip_eq_cicsam_div(s->gTmp, s->uF.p, s->beta.p, 0.5, s->gBc.p, s->g0F.p, s->g0C.p, s->g0F.p),
     "cicsam::div", s->gTmp->err.c_str());
  ck(phase_hip_eq_add(s->gEq, s->gTmp, 1), "operator+", s->gEq->err.c_str());
  const int gIt = solveInto(s->sg, s->gEq, s->gC.p, "gammaEqn");
  ck(phase_hip_field_interpolate_faces(m, 1, s->gC.p, s->gBc.p, s->gF.p), "gamma faces", m->err.c_str());
  ck(phase_hip_field_grad(m, s->gC.p, s->gF.p, s->ggF.p, s->ggC.p), "grad gamma", m->err.c_str());
  mpProperties(s);
  // solveUEqn
  copyD(s->u1C.p, s->u0C.p, 2 * (size_t)N, st);
  copyD(s->u0C.p, s->uC.p, 2 * (size_t)N, st);
  copyD(s->u1F.p, s->u0F.p, 2 * (size_t)F, st);
  copyD(s->u0F.p, s->uF.p, 2 * (size_t)F, st);
  int uIt = 0, pIt = 0;
  for (int corr = 0; corr < s->nCorr; ++corr) {
    ck(phase_hip_field_face_to_cell(m, s->gpF.p, s->rhoC.p, s->rho0F.p, s->gpC.p), "gradP faceToCell", m->err.c_str());
    ck(phase_hip_eq_ddt(s->uEq, dt, s->u0C.p), "fv::ddt", s->uEq->err.c_str());
    ck(phase_hip_eq_scale_cells(s->uEq, s->rhoC.p), "rho *", s->uEq->err.c_str());
    ck(phase_hip_eq_dive(s->tmp, s->u0F.p, s->u1F.p, 0.5, s->uBc.p, s->u0C.p, s->u1C.p, s->u0F.p, s->u1F.p), "fv::dive",
       s->tmp->err.c_str());
    ck(phase_hip_eq_scale_cells(s->tmp, s->rhoC.p), "rho *", s->tmp->err.c_str());
    ck(phase_hip_eq_add(s->uEq, s->tmp, 1), "operator+", s->uEq->err.c_str());
    ck(phase_hip_eq_laplacian(s->tmp, 2, 0., s->muF.p, s->mu0F.p, 0.5, s->uBc.p, s->uF.p, s->u0C.p, s->u0F.p),
       "fv::laplacian", s->tmp->err.c_str());
    // + src::src(fst + sg - gradP) is -= src::src(gradP - (fst + sg))
    hipLaunchKernelGGL(kMomentumSource, dim3(g256(2L * N)), dim3(256), 0, st, 2L * N, s->gpC.p, s->fstC.p, s->sgC.p,
                       s->srcT.p);
    ck(phase_hip_field_src_src(m, 2, s->srcT.p, s->gsrc.p), "src::src", m->err.c_str());
    ck(phase_hip_eq_sub_vector(s->tmp, s->gsrc.p), "operator+", s->tmp->err.c_str());
    ck(phase_hip_eq_add(s->uEq, s->tmp, -1), "operator==", s->uEq->err.c_str());
    hipLaunchKernelGGL(kToBlocks, dim3(g256(N)), dim3(256), 0, st, N, s->uC.p, s->uB.p);
    uIt += solveInto(s->su, s->uEq, s->uB.p, "uEqn");
    hipLaunchKernelGGL(kFromBlocks, dim3(g256(N)), dim3(256), 0, st, N, s->uB.p, s->uC.p);
    hipLaunchKernelGGL(kDtOverRhoUpdate, dim3(g256(2L * N)), dim3(256), 0, st, 2L * N, dt, 1, s->rhoC.p, s->gpC.p,
                       s->uC.p);
    hipLaunchKernelGGL(kMultiphaseFaceVelocity, dim3(g256(F)), dim3(256), 0, st, F, dt, m->dFL.p, m->dFR.p, m->dVol.p,
                       m->dSf.p, s->uBc.p, s->uC.p, s->rhoC.p, s->rhoF.p, s->fstC.p, s->fstF.p, s->sgC.p, s->sgF.p,
                       s->uF.p);
    PH_CHECK(hipGetLastError());
    // solvePEqn
    hipLaunchKernelGGL(kScalarOver, dim3(g256(F)), dim3(256), 0, st, F, dt, s->rhoF.p, s->mob.p);
    ck(phase_hip_eq_laplacian(s->pEq, 3, 0., s->mob.p, nullptr, 1.0, s->pBc.p, s->pF.p, nullptr, nullptr),
       "fv::laplacian", s->pEq->err.c_str());
    ck(phase_hip_field_src_div(m, s->uF.p, s->m.p), "src::div", m->err.c_str());
    ck(phase_hip_eq_sub_vector(s->pEq, s->m.p), "operator==", s->pEq->err.c_str());
    if (s->pin >= 0)
      hipLaunchKernelGGL(kPinRow, dim3(1), dim3(64), 0, st, s->pin, s->pEq->stride, s->pEq->cols.p, s->pEq->vals.p,
                         s->pEq->cnt.p, s->pEq->rhs.p);
    pIt += solveInto(s->sp, s->pEq, s->pC.p, "pEqn");
    ck(phase_hip_field_interpolate_faces(m, 1, s->pC.p, s->pBc.p, s->pF.p), "setBoundaryFaces", m->err.c_str());
    ck(phase_hip_field_grad(m, s->pC.p, s->pF.p, s->gpF.p, s->gpC.p), "grad p", m->err.c_str());
    ck(phase_hip_field_face_to_cell(m, s->gpF.p, s->rhoC.p, s->rhoF.p, s->gpC.p), "gradP faceToCell", m->err.c_str());
    // correctVelocity
    hipLaunchKernelGGL(kDtOverRhoUpdate, dim3(g256(2L * N)), dim3(256), 0, st, 2L * N, dt, 0, s->rhoC.p, s->gpC.p,
                       s->uC.p);
    hipLaunchKernelGGL(kDtOverRhoUpdate, dim3(g256(2L * F)), dim3(256), 0, st, 2L * F, dt, 0, s->rhoF.p, s->gpF.p,
                       s->uF.p);
    PH_CHECK(hipGetLastError());
  }
  if (hist) {
    // FractionalStep::maxDivergenceError on the corrected faces
    ck(phase_hip_field_src_div(m, s->uF.p, s->m.p), "src::div", m->err.c_str());
    hist[0] = gIt;
    hist[1] = uIt;
    hist[2] = pIt;
    hist[3] = mpMaxAbs(s, s->m.p, N);
  }
}

}  // namespace

extern "C" {

int phase_hip_multiphase_destroy(phase_hip_multiphase *s) {
  if (!s) return PHASE_HIP_OK;
  for (phase_hip_equation *e : {s->gEq, s->uEq, s->tmp, s->gTmp, s->pEq}) phase_hip_equation_destroy(e);
  for (phase_hip_solver *v : {s->sg, s->su, s->sp}) phase_hip_solver_destroy(v);
  delete s->st;
  delete s;
  return PHASE_HIP_OK;
}

const char *phase_hip_multiphase_last_error(const phase_hip_multiphase *s) {
  return s ? s->err.c_str() : "null driver";
}

int phase_hip_multiphase_create(phase_hip_multiphase **out, phase_hip_mesh *m, double rho1, double rho2, double mu1,
                                double mu2, double gx, double gy, const int *uBcType, const double *uFace0,
                                const int *pBcType, const double *pFace0, const int *gammaBcType,
                                const double *gammaCell0, int pinCell) {
  if (!out || !m || !uBcType || !uFace0 || !gammaCell0) return PHASE_HIP_EINVAL;
  *out = nullptr;
  phase_hip_multiphase *s = new phase_hip_multiphase();
  s->mesh = m;
  s->N = m->nCells;
  s->F = m->nFaces;
  s->rho1 = rho1;
  s->rho2 = rho2;
  s->mu1 = mu1;
  s->mu2 = mu2;
  s->gx = gx;
  s->gy = gy;
  s->pin = pinCell;
  const int rc = guardedMP(s, [&] {
    if (!(rho1 > 0.) || !(rho2 > 0.) || !(mu1 > 0.) || !(mu2 > 0.))
      fail(PHASE_HIP_EINVAL, "densities and viscosities must be positive");
    if (pinCell < -1 || pinCell >= m->nCells) fail(PHASE_HIP_EINVAL, "pin cell out of range (-1: none)");
    PH_CHECK(hipSetDevice(m->device));
    hipStream_t st = m->stream;
    const int N = s->N, F = s->F;
    ck(phase_hip_equation_create(&s->gEq, m, 1), "gammaEqn", "");
    ck(phase_hip_equation_create(&s->gTmp, m, 1), "gammaEqn", "");
    ck(phase_hip_equation_create(&s->uEq, m, 2), "uEqn", "");
    ck(phase_hip_equation_create(&s->tmp, m, 2), "uEqn", "");
    ck(phase_hip_equation_create(&s->pEq, m, 1), "pEqn", "");
    for (phase_hip_solver **v : {&s->sg, &s->su, &s->sp}) {
      ck(phase_hip_solver_create(v, m->device), "solver", "");
      ck(phase_hip_solver_setup(*v, "tolerance", "1e-8"), "setup", (*v)->err.c_str());
      ck(phase_hip_solver_setup(*v, "maxIters", "5000"), "setup", (*v)->err.c_str());
      // the volume-fraction rows drop the zero upwind coefficients, so their
      // pattern follows the flow: keep one ordering for the union
      ck(phase_hip_solver_setup(*v, "pattern", "union"), "setup", (*v)->err.c_str());
    }
    for (DBuf<double> *b : {&s->gC, &s->g0C, &s->rhoC, &s->muC, &s->pC, &s->m}) b->alloc(N);
    for (DBuf<double> *b : {&s->gF, &s->g0F, &s->beta, &s->rhoF, &s->muF, &s->rho0F, &s->mu0F, &s->pF, &s->mob})
      b->alloc(F);
    for (DBuf<double> *b : {&s->ggC, &s->grC, &s->sgC, &s->uC, &s->u0C, &s->u1C, &s->uB, &s->gpC, &s->gsrc, &s->srcT,
                            &s->fstC})
      b->alloc(2 * (size_t)N);
    for (DBuf<double> *b : {&s->ggF, &s->grF, &s->sgF, &s->uF, &s->u0F, &s->u1F, &s->gpF, &s->fstF})
      b->alloc(2 * (size_t)F);
    s->maxBuf.alloc(1 + kMaxBlocks);
    for (DBuf<double> *b : {&s->uC, &s->pC, &s->gpC, &s->gpF, &s->gF, &s->rhoF, &s->muF, &s->fstC, &s->fstF})
      b->zero(st);
    std::vector<int> ubc(uBcType, uBcType + F), pbc(F, PHASE_HIP_BC_NORMAL_GRADIENT),
        gbc(F, PHASE_HIP_BC_NORMAL_GRADIENT);
    if (pBcType) pbc.assign(pBcType, pBcType + F);
    if (gammaBcType) gbc.assign(gammaBcType, gammaBcType + F);
    s->uBc.upload(ubc, st);
    s->pBc.upload(pbc, st);
    s->gBc.upload(gbc, st);
    PH_CHECK(hipMemcpyAsync(s->uF.p, uFace0, 2 * (size_t)F * sizeof(double), hipMemcpyHostToDevice, st));
    if (pFace0) PH_CHECK(hipMemcpyAsync(s->pF.p, pFace0, (size_t)F * sizeof(double), hipMemcpyHostToDevice, st));
    else s->pF.zero(st);
    PH_CHECK(hipMemcpyAsync(s->gC.p, gammaCell0, (size_t)N * sizeof(double), hipMemcpyHostToDevice, st));
    ck(phase_hip_field_interpolate_faces(m, 2, s->uC.p, s->uBc.p, s->uF.p), "u faces", m->err.c_str());
    copyD(s->u0C.p, s->uC.p, 2 * (size_t)N, st);
    copyD(s->u1C.p, s->uC.p, 2 * (size_t)N, st);
    copyD(s->u0F.p, s->uF.p, 2 * (size_t)F, st);
    copyD(s->u1F.p, s->uF.p, 2 * (size_t)F, st);
    ck(phase_hip_field_interpolate_faces(m, 1, s->pC.p, s->pBc.p, s->pF.p), "p faces", m->err.c_str());
    ck(phase_hip_field_interpolate_faces(m, 1, s->gC.p, s->gBc.p, s->gF.p), "gamma faces", m->err.c_str());
    ck(phase_hip_field_grad(m, s->gC.p, s->gF.p, s->ggF.p, s->ggC.p), "grad gamma", m->err.c_str());
    mpProperties(s);
    PH_CHECK(hipStreamSynchronize(st));
  });
  if (rc != PHASE_HIP_OK) {
    phase_hip_multiphase_destroy(s);
    return rc;
  }
  *out = s;
  return PHASE_HIP_OK;
}

int phase_hip_multiphase_setup(phase_hip_multiphase *s, const char *key, const char *value) {
  return guardedMP(s, [&] {
    if (!key || !value) fail(PHASE_HIP_EINVAL, "null key/value");
    std::string k(key);
    if (k == "numPressureCorrections") {
      s->nCorr = parseCorrections(value);
      return;
    }
    phase_hip_solver *v = nullptr;
    size_t skip = 2;
    if (k.rfind("u.", 0) == 0) v = s->su;
    else if (k.rfind("p.", 0) == 0) v = s->sp;
    else if (k.rfind("gamma.", 0) == 0) v = s->sg, skip = 6;
    else
      fail(PHASE_HIP_EINVAL,
           "keys are numPressureCorrections, gamma.<solver key>, u.<solver key> or p.<solver key>");
    ck(phase_hip_solver_setup(v, k.c_str() + skip, value), "setup", v->err.c_str());
  });
}

int phase_hip_multiphase_advance(phase_hip_multiphase *s, double dt, int nSteps, double *history) {
  return guardedMP(s, [&] {
    if (nSteps < 0 || !(dt > 0.)) fail(PHASE_HIP_EINVAL, "need dt > 0 and nSteps >= 0");
    PH_CHECK(hipSetDevice(s->mesh->device));
    for (int it = 0; it < nSteps; ++it) mpStep(s, dt, history ? history + 4 * (size_t)it : nullptr);
  });
}

int phase_hip_multiphase_surface_tension(phase_hip_multiphase *s, double sigma, double kernelRadius,
                                         int kernelType, const double *contactAngle) {
  return guardedMP(s, [&] {
    if (!(sigma >= 0.)) fail(PHASE_HIP_EINVAL, "sigma must be >= 0");
    PH_CHECK(hipSetDevice(s->mesh->device));
    auto *cm = new phase::CelesteModel();
    try {
      phase::celesteSetup(s->mesh, sigma, kernelRadius, kernelType, contactAngle, *cm);
    } catch (...) {
      delete cm;
      throw;
    }
    delete s->st;
    s->st = cm;
    mpSurfaceForce(s);
    PH_CHECK(hipStreamSynchronize(s->mesh->stream));
  });
}

int phase_hip_multiphase_surface_fields(phase_hip_multiphase *s, double *gammaTilde, double *nCell, double *nFace,
                                        double *kappaCell, double *kappaFace, double *fstFace, double *fstCell) {
  return guardedMP(s, [&] {
    if (!s->st && (gammaTilde || nCell || nFace || kappaCell || kappaFace))
      fail(PHASE_HIP_EINVAL, "surface tension is not turned on");
    hipStream_t st = s->mesh->stream;
    const size_t N = s->N, F = s->F, d = sizeof(double);
    auto get = [&](double *h, const double *dv, size_t n) {
      if (h) PH_CHECK(hipMemcpyAsync(h, dv, n * d, hipMemcpyDeviceToHost, st));
    };
    if (s->st) {
      get(gammaTilde, s->st->gt.p, N);
      get(nCell, s->st->nC.p, 2 * N);
      get(nFace, s->st->nF.p, 2 * F);
      get(kappaCell, s->st->kC.p, N);
      get(kappaFace, s->st->kF.p, F);
    }
    get(fstFace, s->fstF.p, 2 * F);
    get(fstCell, s->fstC.p, 2 * N);
    PH_CHECK(hipStreamSynchronize(st));
  });
}

int phase_hip_multiphase_fields(phase_hip_multiphase *s, double *uCell, double *pCell, double *uFace,
                                double *gammaCell) {
  return guardedMP(s, [&] {
    hipStream_t st = s->mesh->stream;
    const size_t N = s->N, F = s->F;
    if (uCell) PH_CHECK(hipMemcpyAsync(uCell, s->uC.p, 2 * N * sizeof(double), hipMemcpyDeviceToHost, st));
    if (pCell) PH_CHECK(hipMemcpyAsync(pCell, s->pC.p, N * sizeof(double), hipMemcpyDeviceToHost, st));
    if (uFace) PH_CHECK(hipMemcpyAsync(uFace, s->uF.p, 2 * F * sizeof(double), hipMemcpyDeviceToHost, st));
    if (gammaCell) PH_CHECK(hipMemcpyAsync(gammaCell, s->gC.p, N * sizeof(double), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipStreamSynchronize(st));
  });
}

}  // extern "C"
