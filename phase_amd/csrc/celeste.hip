// CELESTE surface tension (src/2D/Unstructured/FiniteVolume/Multiphase/Celeste.cpp:9-109,
// CelesteStencil.cpp:5-137, SurfaceTensionForce.cpp:48-78,
// SurfaceTensionForceSmoothingKernel.cpp:3-59).
//
// Set-up, host, once per mesh:
//   smoothing kernel of each cell: the cells whose centroid lies strictly
//     inside the circle of radius eps around it (Circle::isInside), in
//     ascending id, with K(x_k - x) and A = 1 / sum K V;
//   stencil of each cell: face neighbours, diagonal (node) neighbours, then
//     the neighbours' boundary faces (cells touching the boundary only) and
//     the cell's own boundary faces; least-squares matrix rows
//     (dx^2/2, dy^2/2, dx dy, dx, dy) / s (s = |dx|^2 weighted, 1 unweighted);
//     rows 3 and 4 of its pseudo-inverse by Householder QR (what LAPACK dgels
//     computes for pseudoInverse, Math/Matrix.cpp:231-237).
// Per step, device:
//   kSmooth:        gammaTilde = A sum (gamma_k K_k) V_k; boundary faces = cell
//   kGradNormal:    gradGammaTilde (weighted stencil); n = -grad / |grad| or 0
//   kWallNormal:    boundary-face n from the contact angle
//   kCurvature:     kappa = div n (unweighted stencil) where n != 0 on the cell,
//                   its face and its diagonal neighbours
//   kFaceForce:     kappa_f (volume-weighted / one-sided), fst_f = sigma kappa_f gradGamma_f
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "celeste.h"

namespace phase {
namespace {

inline int g256c(long n) { return (int)std::max<long>(1, std::min<long>((n + 255) / 256, 8192)); }

constexpr double kEps = 1e-8;  // SurfaceTensionForce.h: eps_ = 1e-8

template <class T>
std::vector<T> download(const DBuf<T> &d, size_t n, hipStream_t st) {
  std::vector<T> h(n);
  if (n) PH_CHECK(hipMemcpyAsync(h.data(), d.p, n * sizeof(T), hipMemcpyDeviceToHost, st));
  PH_CHECK(hipStreamSynchronize(st));
  return h;
}

// SmoothingKernel::kernel, SurfaceTensionForceSmoothingKernel.cpp:45-58
double kernelValue(double dx, double dy, double eps, int type) {
  if (type == PHASE_HIP_KERNEL_PESKIN) {
    auto kcos = [eps](double x) { return x < eps ? eps * (1. + std::cos(M_PI * x / eps)) : 0.; };
    return kcos(dx) * kcos(dy);
  }
  const double r = std::sqrt(dx * dx + dy * dy);
  const double p = type == PHASE_HIP_KERNEL_POW6 ? 3. : 4.;
  return r < eps ? std::pow(std::pow(eps, 2) - std::pow(r, 2), p) : 0.;
}

// Rows 3 and 4 of pinv(A) for a k x 5 row-major A of full column rank:
// Householder QR, then R^-1 Q^T e_t for every t.  Returns false when a
// column is (numerically) dependent.
bool pinvRows34(int k, std::vector<double> a, double *row3, double *row4) {
  if (k < 5) return false;
  std::vector<std::vector<double>> vs(5);
  double vn[5];
  double rmax = 0.;
  for (int j = 0; j < 5; ++j) {
    double nrm = 0.;
    for (int i = j; i < k; ++i) nrm += a[i * 5 + j] * a[i * 5 + j];
    nrm = std::sqrt(nrm);
    std::vector<double> v(k - j);
    for (int i = j; i < k; ++i) v[i - j] = a[i * 5 + j];
    const double alpha = a[j * 5 + j] > 0. ? -nrm : nrm;
    v[0] -= alpha;
    double vv = 0.;
    for (double x : v) vv += x * x;
    if (vv > 0.) {
      for (int c = j; c < 5; ++c) {
        double s = 0.;
        for (int i = j; i < k; ++i) s += v[i - j] * a[i * 5 + c];
        const double f = 2. * s / vv;
        for (int i = j; i < k; ++i) a[i * 5 + c] -= f * v[i - j];
      }
    }
    vs[j] = std::move(v);
    vn[j] = vv;
    rmax = std::max(rmax, std::fabs(a[j * 5 + j]));
  }
  for (int j = 0; j < 5; ++j)
    if (!(std::fabs(a[j * 5 + j]) > 1e-13 * rmax)) return false;
  std::vector<double> y(k);
  for (int t = 0; t < k; ++t) {
    std::fill(y.begin(), y.end(), 0.);
    y[t] = 1.;
    for (int j = 0; j < 5; ++j) {
      if (vn[j] == 0.) continue;
      double s = 0.;
      for (int i = j; i < k; ++i) s += vs[j][i - j] * y[i];
      const double f = 2. * s / vn[j];
      for (int i = j; i < k; ++i) y[i] -= f * vs[j][i - j];
    }
    double z[5];
    for (int i = 4; i >= 0; --i) {
      double acc = y[i];
      for (int c = i + 1; c < 5; ++c) acc -= a[i * 5 + c] * z[c];
      z[i] = acc / a[i * 5 + i];
    }
    row3[t] = z[3];
    row4[t] = z[4];
  }
  return true;
}

// ---------------------------------------------------------------- kernels
__global__ void kSmooth(int n, const int *__restrict__ kptr, const int *__restrict__ kcol,
                        const double *__restrict__ kw, const double *__restrict__ A, const double *__restrict__ vol,
                        const double *__restrict__ gamma, double *__restrict__ gt) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    double acc = 0.;
    for (int e = kptr[c]; e < kptr[c + 1]; ++e) {
      const int j = kcol[e];
      acc += gamma[j] * kw[e] * vol[j];
    }
    gt[c] = A[c] * acc;
  }
}

// smoothGammaField: fill(0) then setBoundaryFaces (zero gradient)
__global__ void kTildeFaces(int nFaces, const int *__restrict__ fl, const int *__restrict__ fr,
                            const double *__restrict__ gt, double *__restrict__ gtf) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nFaces; f += gridDim.x * blockDim.x)
    gtf[f] = fr[f] < 0 ? gt[fl[f]] : 0.;
}

// Stencil::grad (weighted) and computeInterfaceNormals' cell part
__global__ void kGradNormal(int n, const int *__restrict__ sptr, const int *__restrict__ sidx,
                            const double *__restrict__ ss, const double *__restrict__ pw,
                            const double *__restrict__ gt, const double *__restrict__ gtf, double *__restrict__ ggt,
                            double *__restrict__ nC) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const double p0 = gt[c];
    double gx = 0., gy = 0.;
    for (int e = sptr[c]; e < sptr[c + 1]; ++e) {
      const int k = sidx[e];
      const double v = k >= 0 ? gt[k] : gtf[-k - 1];
      const double b = (v - p0) / ss[e];
      gx += pw[2 * e] * b;
      gy += pw[2 * e + 1] * b;
    }
    ggt[2 * c] = gx;
    ggt[2 * c + 1] = gy;
    const double m2 = gx * gx + gy * gy;
    if (m2 >= kEps * kEps) {
      const double mag = sqrt(m2);
      nC[2 * c] = -(gx / mag);
      nC[2 * c + 1] = -(gy / mag);
    } else {
      nC[2 * c] = 0.;
      nC[2 * c + 1] = 0.;
    }
  }
}

// computeInterfaceNormals' boundary faces (SurfaceTensionForce.cpp:59-77)
__global__ void kWallNormal(int nFaces, const int *__restrict__ fl, const int *__restrict__ fr,
                            const double *__restrict__ fc, const double *__restrict__ cc,
                            const double *__restrict__ sf, const double *__restrict__ cosT,
                            const double *__restrict__ sinT, const double *__restrict__ nC, double *__restrict__ nF) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nFaces; f += gridDim.x * blockDim.x) {
    if (fr[f] >= 0) continue;
    const int l = fl[f];
    const double nx = nC[2 * l], ny = nC[2 * l + 1];
    if (nx * nx + ny * ny == 0.) {
      nF[2 * f] = 0.;
      nF[2 * f + 1] = 0.;
      continue;
    }
    // ns = -outwardNorm(lCell centroid).unitVec()
    double ox = sf[2 * f], oy = sf[2 * f + 1];
    const double dx = fc[2 * f] - cc[2 * l], dy = fc[2 * f + 1] - cc[2 * l + 1];
    if (!(dx * ox + dy * oy > 0.)) {
      ox = -ox;
      oy = -oy;
    }
    const double om = sqrt(ox * ox + oy * oy);
    const double sx = -(ox / om), sy = -(oy / om);
    // ts = n.tangentialComponent(ns).unitVec()
    const double t = (nx * sx + ny * sy) / (sx * sx + sy * sy);
    const double tx = nx - t * sx, ty = ny - t * sy;
    const double tm = sqrt(tx * tx + ty * ty);
    const double ux = tx / tm, uy = ty / tm;
    if (isnan(ux) || isnan(uy)) {
      nF[2 * f] = nx;
      nF[2 * f + 1] = ny;
    } else {
      nF[2 * f] = sx * cosT[f] + ux * sinT[f];
      nF[2 * f + 1] = sy * cosT[f] + uy * sinT[f];
    }
  }
}

__device__ __forceinline__ bool nonzero2(const double *v, int i) {
  return v[2 * i] * v[2 * i] + v[2 * i + 1] * v[2 * i + 1] != 0.;
}

// Celeste::computeCurvature's cell part (Celeste.cpp:53-79)
__global__ void kCurvature(int n, const int *__restrict__ iptr, const int *__restrict__ inb,
                           const int *__restrict__ dptr, const int *__restrict__ dcol, const int *__restrict__ sptr,
                           const int *__restrict__ sidx, const double *__restrict__ pu,
                           const double *__restrict__ nC, const double *__restrict__ nF, double *__restrict__ kC) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    bool valid = nonzero2(nC, c);
    for (int e = iptr[c]; valid && e < iptr[c + 1]; ++e) valid = nonzero2(nC, inb[e]);
    for (int e = dptr[c]; valid && e < dptr[c + 1]; ++e) valid = nonzero2(nC, dcol[e]);
    if (!valid) {
      kC[c] = 0.;
      continue;
    }
    const double cx = nC[2 * c], cy = nC[2 * c + 1];
    double ax = 0., ay = 0.;
    for (int e = sptr[c]; e < sptr[c + 1]; ++e) {
      const int k = sidx[e];
      const double vx = k >= 0 ? nC[2 * k] : nF[2 * (-k - 1)];
      const double vy = k >= 0 ? nC[2 * k + 1] : nF[2 * (-k - 1) + 1];
      ax += pu[2 * e] * (vx - cx);
      ay += pu[2 * e + 1] * (vy - cy);
    }
    kC[c] = ax + ay;
  }
}

// Celeste::computeCurvature's face part (Celeste.cpp:81-98) and
// computeFaceInterfaceForces' fst_f = sigma kappa_f gradGamma_f (:17-19)
__global__ void kFaceForce(int nFaces, double sigma, const int *__restrict__ fl, const int *__restrict__ fr,
                           const double *__restrict__ vol, const double *__restrict__ nC,
                           const double *__restrict__ kC, const double *__restrict__ ggF, double *__restrict__ kF,
                           double *__restrict__ fst) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nFaces; f += gridDim.x * blockDim.x) {
    const int a = fl[f], b = fr[f];
    const bool na = nonzero2(nC, a);
    if (b >= 0) {
      const bool nb = nonzero2(nC, b);
      if (na && nb) {
        const double v1 = vol[b], v2 = vol[a];
        const double g = v1 / (v1 + v2);
        kF[f] = g * kC[a] + (1. - g) * kC[b];
      } else if (na) {
        kF[f] = kC[a];
      } else if (nb) {
        kF[f] = kC[b];
      } else {
        kF[f] = 0.;
      }
    } else if (na) {
      kF[f] = kC[a];
    }
    const double s = sigma * kF[f];
    fst[2 * f] = s * ggF[2 * f];
    fst[2 * f + 1] = s * ggF[2 * f + 1];
  }
}

}  // namespace

void celesteSetup(phase_hip_mesh *m, double sigma, double kernelWidth, int kernelType, const double *contactAngle,
                  CelesteModel &cm) {
  if (!(kernelWidth > 0.)) fail(PHASE_HIP_EINVAL, "smoothing kernel radius must be positive");
  if (kernelType != PHASE_HIP_KERNEL_PESKIN && kernelType != PHASE_HIP_KERNEL_POW6 &&
      kernelType != PHASE_HIP_KERNEL_POW8)
    fail(PHASE_HIP_EINVAL, "kernel type is PESKIN, POW6 or POW8");
  hipStream_t st = m->stream;
  const int N = m->nCells, F = m->nFaces;
  cm.N = N;
  cm.F = F;
  cm.sigma = sigma;
  const std::vector<double> cc = download(m->dCC, 2 * (size_t)N, st), fc = download(m->dFC, 2 * (size_t)F, st),
                            vol = download(m->dVol, N, st);
  const std::vector<int> cptr = download(m->dCptr, N + 1, st);
  const std::vector<int> cind = download(m->dCind, cptr[N], st);
  // Cell::diagonals, FiniteVolumeGrid2D.cpp:423-430
  std::vector<std::vector<int>> nodeCells(m->nNodes);
  for (int c = 0; c < N; ++c)
    for (int k = cptr[c]; k < cptr[c + 1]; ++k) nodeCells[cind[k]].push_back(c);
  std::vector<int> dptr(N + 1, 0), dcol;
  for (int c = 0; c < N; ++c) {
    for (int k = cptr[c]; k < cptr[c + 1]; ++k)
      for (int kc : nodeCells[cind[k]]) {
        if (kc == c) continue;
        bool nb = false;
        for (int e = m->iptr[c]; e < m->iptr[c + 1] && !nb; ++e) nb = m->inb[e] == kc;
        if (!nb) dcol.push_back(kc);
      }
    dptr[c + 1] = (int)dcol.size();
  }
  // stencils, CelesteStencil.cpp:9-31, 112-137
  std::vector<int> sptr(N + 1, 0), sidx;
  std::vector<double> ss, pw, pu;
  std::vector<double> rowsW, rowsU;
  for (int c = 0; c < N; ++c) {
    std::vector<int> cells, faces;
    const bool hasBd = m->bptr[c + 1] > m->bptr[c];
    for (int e = m->iptr[c]; e < m->iptr[c + 1]; ++e) {
      const int nb = m->inb[e];
      cells.push_back(nb);
      if (hasBd)
        for (int q = m->bptr[nb]; q < m->bptr[nb + 1]; ++q) faces.push_back(m->bface[q]);
    }
    for (int e = dptr[c]; e < dptr[c + 1]; ++e) cells.push_back(dcol[e]);
    for (int q = m->bptr[c]; q < m->bptr[c + 1]; ++q) faces.push_back(m->bface[q]);
    const int k = (int)(cells.size() + faces.size());
    rowsW.assign(5 * (size_t)k, 0.);
    rowsU.assign(5 * (size_t)k, 0.);
    const double x0 = cc[2 * c], y0 = cc[2 * c + 1];
    int i = 0;
    for (int kc : cells) {
      const double rx = cc[2 * kc] - x0, ry = cc[2 * kc + 1] - y0;
      const double s = rx * rx + ry * ry;
      const double w[5] = {rx * rx / 2. / s, ry * ry / 2. / s, rx * ry / s, rx / s, ry / s};
      const double u[5] = {rx * rx / 2. / 1., ry * ry / 2. / 1., rx * ry / 1., rx / 1., ry / 1.};
      std::copy(w, w + 5, &rowsW[5 * (size_t)i]);
      std::copy(u, u + 5, &rowsU[5 * (size_t)i]);
      sidx.push_back(kc);
      ss.push_back(s);
      ++i;
    }
    for (int f : faces) {
      const double rx = fc[2 * f] - x0, ry = fc[2 * f + 1] - y0;
      const double s = rx * rx + ry * ry;
      const double w[5] = {rx * rx / (2. * s), ry * ry / (2. * s), rx * ry / s, rx / s, ry / s};
      const double u[5] = {rx * rx / (2. * 1.), ry * ry / (2. * 1.), rx * ry / 1., rx / 1., ry / 1.};
      std::copy(w, w + 5, &rowsW[5 * (size_t)i]);
      std::copy(u, u + 5, &rowsU[5 * (size_t)i]);
      sidx.push_back(-f - 1);
      ss.push_back(s);
      ++i;
    }
    std::vector<double> r3(k), r4(k), q3(k), q4(k);
    if (!pinvRows34(k, rowsW, r3.data(), r4.data()) || !pinvRows34(k, rowsU, q3.data(), q4.data()))
      fail(PHASE_HIP_EINVAL, "surface-tension stencil of cell " + std::to_string(c) + " is rank deficient");
    for (int t = 0; t < k; ++t) {
      pw.push_back(r3[t]);
      pw.push_back(r4[t]);
      pu.push_back(q3[t]);
      pu.push_back(q4[t]);
    }
    sptr[c + 1] = (int)sidx.size();
  }
  // smoothing kernels over a uniform bin grid of cell centroids
  double xmin = cc[0], xmax = cc[0], ymin = cc[1], ymax = cc[1];
  for (int c = 0; c < N; ++c) {
    xmin = std::min(xmin, cc[2 * c]);
    xmax = std::max(xmax, cc[2 * c]);
    ymin = std::min(ymin, cc[2 * c + 1]);
    ymax = std::max(ymax, cc[2 * c + 1]);
  }
  const double eps = kernelWidth;
  const long nbx = std::max<long>(1, std::min<long>(4096, (long)((xmax - xmin) / eps) + 1));
  const long nby = std::max<long>(1, std::min<long>(4096, (long)((ymax - ymin) / eps) + 1));
  const double bx = (xmax - xmin) / nbx + 1e-300, by = (ymax - ymin) / nby + 1e-300;
  auto binOf = [&](double v, double lo, double w, long nb) {
    return std::min<long>(nb - 1, std::max<long>(0, (long)((v - lo) / w)));
  };
  std::vector<int> bptr(nbx * nby + 1, 0), bcell(N);
  for (int c = 0; c < N; ++c) ++bptr[binOf(cc[2 * c + 1], ymin, by, nby) * nbx + binOf(cc[2 * c], xmin, bx, nbx) + 1];
  for (long b = 0; b < nbx * nby; ++b) bptr[b + 1] += bptr[b];
  {
    std::vector<int> fill(bptr.begin(), bptr.end() - 1);
    for (int c = 0; c < N; ++c)
      bcell[fill[binOf(cc[2 * c + 1], ymin, by, nby) * nbx + binOf(cc[2 * c], xmin, bx, nbx)]++] = c;
  }
  const long rx = (long)std::ceil(eps / bx), ry = (long)std::ceil(eps / by);
  std::vector<int> kptr(N + 1, 0), kcol, cand;
  std::vector<double> kw, A(N);
  for (int c = 0; c < N; ++c) {
    const double x0 = cc[2 * c], y0 = cc[2 * c + 1];
    const long ix = binOf(x0, xmin, bx, nbx), iy = binOf(y0, ymin, by, nby);
    cand.clear();
    for (long j = std::max<long>(0, iy - ry); j <= std::min<long>(nby - 1, iy + ry); ++j)
      for (long i = std::max<long>(0, ix - rx); i <= std::min<long>(nbx - 1, ix + rx); ++i)
        for (int q = bptr[j * nbx + i]; q < bptr[j * nbx + i + 1]; ++q) {
          const int k = bcell[q];
          const double dx = cc[2 * k] - x0, dy = cc[2 * k + 1] - y0;
          if (dx * dx + dy * dy < eps * eps) cand.push_back(k);  // Circle::isInside
        }
    std::sort(cand.begin(), cand.end());
    double a = 0.;
    for (int k : cand) {
      const double w = kernelValue(cc[2 * k] - x0, cc[2 * k + 1] - y0, eps, kernelType);
      kcol.push_back(k);
      kw.push_back(w);
      a += w * vol[k];
    }
    A[c] = 1. / a;
    kptr[c + 1] = (int)kcol.size();
  }
  std::vector<double> cosT(F, std::cos(M_PI_2)), sinT(F, std::sin(M_PI_2));
  if (contactAngle)
    for (int f = 0; f < F; ++f) {
      cosT[f] = std::cos(contactAngle[f]);
      sinT[f] = std::sin(contactAngle[f]);
    }
  cm.kptr.upload(kptr, st);
  cm.kcol.upload(kcol, st);
  cm.kw.upload(kw, st);
  cm.A.upload(A, st);
  cm.sptr.upload(sptr, st);
  cm.sidx.upload(sidx, st);
  cm.ss.upload(ss, st);
  cm.pw.upload(pw, st);
  cm.pu.upload(pu, st);
  cm.dptr.upload(dptr, st);
  cm.dcol.upload(dcol, st);
  cm.cosT.upload(cosT, st);
  cm.sinT.upload(sinT, st);
  for (DBuf<double> *b : {&cm.gt, &cm.kC}) b->alloc(N);
  for (DBuf<double> *b : {&cm.gtf, &cm.kF}) b->alloc(F);
  for (DBuf<double> *b : {&cm.ggt, &cm.nC}) b->alloc(2 * (size_t)N);
  cm.nF.alloc(2 * (size_t)F);
  for (DBuf<double> *b : {&cm.gt, &cm.kC, &cm.gtf, &cm.kF, &cm.ggt, &cm.nC, &cm.nF}) b->zero(st);
  PH_CHECK(hipStreamSynchronize(st));
}

void celesteFaceForces(phase_hip_mesh *m, CelesteModel &cm, const double *gammaCell, const double *gradGammaFace,
                       double *fstFace) {
  hipStream_t st = m->stream;
  const int N = cm.N, F = cm.F;
  hipLaunchKernelGGL(kSmooth, dim3(g256c(N)), dim3(256), 0, st, N, cm.kptr.p, cm.kcol.p, cm.kw.p, cm.A.p, m->dVol.p,
                     gammaCell, cm.gt.p);
  hipLaunchKernelGGL(kTildeFaces, dim3(g256c(F)), dim3(256), 0, st, F, m->dFL.p, m->dFR.p, cm.gt.p, cm.gtf.p);
  hipLaunchKernelGGL(kGradNormal, dim3(g256c(N)), dim3(256), 0, st, N, cm.sptr.p, cm.sidx.p, cm.ss.p, cm.pw.p,
                     cm.gt.p, cm.gtf.p, cm.ggt.p, cm.nC.p);
  hipLaunchKernelGGL(kWallNormal, dim3(g256c(F)), dim3(256), 0, st, F, m->dFL.p, m->dFR.p, m->dFC.p, m->dCC.p,
                     m->dSf.p, cm.cosT.p, cm.sinT.p, cm.nC.p, cm.nF.p);
  hipLaunchKernelGGL(kCurvature, dim3(g256c(N)), dim3(256), 0, st, N, m->dIptr.p, m->dInb.p, cm.dptr.p, cm.dcol.p,
                     cm.sptr.p, cm.sidx.p, cm.pu.p, cm.nC.p, cm.nF.p, cm.kC.p);
  hipLaunchKernelGGL(kFaceForce, dim3(g256c(F)), dim3(256), 0, st, F, cm.sigma, m->dFL.p, m->dFR.p, m->dVol.p,
                     cm.nC.p, cm.kC.p, gradGammaFace, cm.kF.p, fstFace);
  PH_CHECK(hipGetLastError());
}

}  // namespace phase
