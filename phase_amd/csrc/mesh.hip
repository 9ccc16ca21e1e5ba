// Unstructured 2D mesh on the device and the fv::laplacian assembly kernel.
//
// Topology follows FiniteVolumeGrid2D::init(nodes, cptr, cind)
// (src/2D/Unstructured/FiniteVolumeGrid2D/FiniteVolumeGrid2D.cpp:20-35,
// createCell :84-114, init() :396-418): a face per distinct node pair, numbered
// in first-encounter order over the cells' edges; the creating cell is lCell.
// Each cell lists its interior links, then its boundary links, both by
// ascending face id.  That order fixes the reference's per-row summation order
// and first-touch column order, so the device assembly can reproduce the
// reference's CSR bit for bit.
//
// Geometry is computed on the device with the reference's formulas and no
// contraction (-ffp-contract=off): shoelace area and Bashein-Detmer centroid
// (the Boost.Geometry routines behind Polygon.cpp:218-231), face centroid and
// unnormalised normal (Face.cpp:9-18), the outward sign test (Face.cpp:48-50)
// and the per-link vectors (CellLink.cpp:8, InteriorLink.cpp:9,
// BoundaryLink.cpp:8).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <unordered_map>
#include <vector>

#include "mesh.h"
#include "solver.h"
#include "solver_kernels.h"

using namespace phase;


namespace phase {

// Assembly binding of a solver to a mesh: where each link's coefficient lands.
struct MeshBinding {
  std::vector<int> rowPtr, colInd;   // compact pattern, original numbering
  std::vector<int> islot;            // per interior link: compact slot of (P, nb)
  std::vector<uint8_t> ifirst;       // 1 if the link is the first to touch its slot
  std::vector<int> dslot;            // per cell: compact slot of (P, P)
  DBuf<int> dIslot, dDslot;
  DBuf<uint8_t> dIfirst;
  DBuf<double> dRhs;                 // stored rhs (A x + rhs = 0), original numbering
  DBuf<double> dStage[6];            // uploaded operator inputs
  DBuf<int> dBc;
  DBuf<int> dErr;
  std::vector<int> lastBc;           // boundary types of the last assembly (host copy)
  bool bcOnHost = false;
};

void solverFreeBinding(phase_hip_solver *s) {
  delete s->binding;
  s->binding = nullptr;
  s->boundMesh = nullptr;
  s->lastVariant = -1;
}

}  // namespace phase

namespace {


__global__ void kCellGeom(int nCells, const int *__restrict__ cptr, const int *__restrict__ cind,
                          const double *__restrict__ xy, double *__restrict__ vol, double *__restrict__ cc) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nCells; c += gridDim.x * blockDim.x) {
    const int b = cptr[c], nv = cptr[c + 1] - b;
    double a2 = 0., sx = 0., sy = 0.;
    for (int k = 0; k < nv; ++k) {
      const int i = cind[b + k], j = cind[b + (k + 1 == nv ? 0 : k + 1)];
      const double px = xy[2 * i], py = xy[2 * i + 1], qx = xy[2 * j], qy = xy[2 * j + 1];
      const double ai = px * qy - qx * py;
      a2 += ai;
      sx += (px + qx) * ai;
      sy += (py + qy) * ai;
    }
    if (a2 < 0.) {
      a2 = -a2;
      sx = -sx;
      sy = -sy;
    }
    vol[c] = a2 / 2.;
    cc[2 * c] = sx / (3. * a2);
    cc[2 * c + 1] = sy / (3. * a2);
  }
}

__global__ void kFaceGeom(int nFaces, const int *__restrict__ n1, const int *__restrict__ n2,
                          const double *__restrict__ xy, double *__restrict__ fc, double *__restrict__ sf) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nFaces; f += gridDim.x * blockDim.x) {
    const int a = n1[f], b = n2[f];
    const double ax = xy[2 * a], ay = xy[2 * a + 1], bx = xy[2 * b], by = xy[2 * b + 1];
    fc[2 * f] = 0.5 * (ax + bx);
    fc[2 * f + 1] = 0.5 * (ay + by);
    const double tx = bx - ax, ty = by - ay;
    sf[2 * f] = ty;   // Vector2D::normalVec() = (y, -x)
    sf[2 * f + 1] = -tx;
  }
}

// Per link: dot(r, Sf_out) and |r|^2 with r = c_nb - c_P (interior) or
// c_f - c_P (boundary), exactly as fv::laplacian evaluates them.
__global__ void kLinkGeom(int nCells, const double *__restrict__ cc, const double *__restrict__ fc,
                          const double *__restrict__ sf, const int *__restrict__ iptr,
                          const int *__restrict__ iface, const int *__restrict__ inb,
                          const int *__restrict__ bptr, const int *__restrict__ bface,
                          double *__restrict__ idot, double *__restrict__ ims, double *__restrict__ bdot,
                          double *__restrict__ bms) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nCells; c += gridDim.x * blockDim.x) {
    const Dv2 p{cc[2 * c], cc[2 * c + 1]};
    for (int k = iptr[c]; k < iptr[c + 1]; ++k) {
      const int f = iface[k], nb = inb[k];
      const Dv2 s = outward(Dv2{fc[2 * f], fc[2 * f + 1]}, Dv2{sf[2 * f], sf[2 * f + 1]}, p);
      const double rx = cc[2 * nb] - p.x, ry = cc[2 * nb + 1] - p.y;
      idot[k] = rx * s.x + ry * s.y;
      ims[k] = rx * rx + ry * ry;
    }
    for (int k = bptr[c]; k < bptr[c + 1]; ++k) {
      const int f = bface[k];
      const Dv2 fcen{fc[2 * f], fc[2 * f + 1]};
      const Dv2 s = outward(fcen, Dv2{sf[2 * f], sf[2 * f + 1]}, p);
      const double rx = fcen.x - p.x, ry = fcen.y - p.y;
      bdot[k] = rx * s.x + ry * s.y;
      bms[k] = rx * rx + ry * ry;
    }
  }
}

// fv::laplacian, thread per cell, replaying the reference's per-cell face
// loop (Laplacian.h:7-167): coefficient (gamma * dot) / |r|^2, off-diagonal
// slots first-touch assigned, diagonal and rhs accumulated in link order.
__global__ void kAssembleLaplacian(int nCells, int variant, double gamma, const double *__restrict__ gammaF,
                                   const double *__restrict__ gamma0F, double theta,
                                   const int *__restrict__ bcType, const double *__restrict__ phiF,
                                   const double *__restrict__ phi0C, const double *__restrict__ phi0F,
                                   const int *__restrict__ iptr, const int *__restrict__ iface,
                                   const int *__restrict__ inb, const double *__restrict__ idot,
                                   const double *__restrict__ ims, const int *__restrict__ islot,
                                   const uint8_t *__restrict__ ifirst, const int *__restrict__ bptr,
                                   const int *__restrict__ bface, const double *__restrict__ bdot,
                                   const double *__restrict__ bms, const int *__restrict__ dslot,
                                   double *__restrict__ vals, double *__restrict__ rhs, int *err) {
  const bool hasTheta = (variant == 0 || variant == 2);
  const bool field = (variant >= 2);
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nCells; c += gridDim.x * blockDim.x) {
    double diag = 0., r = 0.;
    bool touched = false;
    for (int k = iptr[c]; k < iptr[c + 1]; ++k) {
      const int f = iface[k];
      const double g = field ? gammaF[f] : gamma;
      const double coeff = g * idot[k] / ims[k];
      double off, dg;
      if (!hasTheta) {
        off = coeff;
        dg = -coeff;
      } else {
        off = theta * coeff;
        dg = theta * -coeff;
        const double coeff0 = field ? gamma0F[f] * idot[k] / ims[k] : coeff;
        r += (1. - theta) * coeff0 * (phi0C[inb[k]] - phi0C[c]);
      }
      const int sl = islot[k];
      vals[sl] = ifirst[k] ? off : vals[sl] + off;
      diag = touched ? diag + dg : dg;
      touched = true;
    }
    for (int k = bptr[c]; k < bptr[c + 1]; ++k) {
      const int f = bface[k];
      const int bt = bcType ? bcType[f] : PHASE_HIP_BC_NORMAL_GRADIENT;
      if (bt == PHASE_HIP_BC_FIXED) {
        if (!phiF || (hasTheta && !phi0F)) {
          atomicOr(err, 2);
          continue;
        }
        const double g = field ? gammaF[f] : gamma;
        const double coeff = g * bdot[k] / bms[k];
        double dg;
        if (!hasTheta) {
          dg = -coeff;
          r += coeff * phiF[f];
        } else {
          dg = theta * -coeff;
          r += theta * coeff * phiF[f];
          const double coeff0 = field ? gamma0F[f] * bdot[k] / bms[k] : coeff;
          r += (1. - theta) * coeff0 * (phi0F[f] - phi0C[c]);
        }
        diag = touched ? diag + dg : dg;
        touched = true;
      } else if (bt != PHASE_HIP_BC_NORMAL_GRADIENT && bt != PHASE_HIP_BC_SYMMETRY) {
        atomicOr(err, 1);
      }
    }
    vals[dslot[c]] = touched ? diag : 0.;
    rhs[c] = r;
  }
}

// b[k] = -rhs[perm[k]]
__global__ void kRhsToB(int n, const int *__restrict__ perm, const double *__restrict__ rhs, double *__restrict__ b) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) b[k] = -rhs[perm[k]];
}

// stored rhs -= v  <=>  b += v
__global__ void kSubRhs(int n, const int *__restrict__ perm, const double *__restrict__ v, double *__restrict__ rhs,
                        double *__restrict__ b) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int i = perm[k];
    rhs[i] -= v[i];
    b[k] = -rhs[i];
  }
}

inline int grid1(long n) { return (int)std::max<long>(1, std::min<long>((n + 255) / 256, 8192)); }

template <class F>
int guardedMesh(phase_hip_mesh *m, F &&f) {
  if (!m) return PHASE_HIP_EINVAL;
  try {
    f();
    return PHASE_HIP_OK;
  } catch (const HipError &e) {
    m->err = e.msg;
    return e.code;
  } catch (const std::exception &e) {
    m->err = e.what();
    return PHASE_HIP_EINVAL;
  }
}

template <class F>
int guardedSolver(phase_hip_solver *s, F &&f) {
  if (!s) return PHASE_HIP_EINVAL;
  try {
    f();
    return PHASE_HIP_OK;
  } catch (const HipError &e) {
    s->err = e.msg;
    return e.code;
  } catch (const std::exception &e) {
    s->err = e.what();
    return PHASE_HIP_EINVAL;
  }
}

template <class T>
void uploadOpt(DBuf<T> &d, const T *h, size_t n, hipStream_t st) {
  if (!h) return;
  d.alloc(n);
  PH_CHECK(hipMemcpyAsync(d.p, h, n * sizeof(T), hipMemcpyHostToDevice, st));
}

// Compact pattern of the cell graph in the first-touch column order of the
// scalar-gamma laplacian (Laplacian.h:57-58: (P,nb) then (P,P) per link).
void buildBinding(phase_hip_mesh *m, MeshBinding &b) {
  const int N = m->nCells;
  b.rowPtr.assign(N + 1, 0);
  b.colInd.clear();
  b.islot.assign(m->iface.size(), -1);
  b.ifirst.assign(m->iface.size(), 0);
  b.dslot.assign(N, -1);
  std::vector<int> cols;
  for (int c = 0; c < N; ++c) {
    cols.clear();
    const int base = (int)b.colInd.size();
    for (int k = m->iptr[c]; k < m->iptr[c + 1]; ++k) {
      const int nb = m->inb[k];
      auto it = std::find(cols.begin(), cols.end(), nb);
      if (it == cols.end()) {
        b.islot[k] = base + (int)cols.size();
        b.ifirst[k] = 1;
        cols.push_back(nb);
      } else {
        b.islot[k] = base + (int)(it - cols.begin());
      }
      if (std::find(cols.begin(), cols.end(), c) == cols.end()) {
        b.dslot[c] = base + (int)cols.size();
        cols.push_back(c);
      }
    }
    if (b.dslot[c] < 0) {
      b.dslot[c] = base + (int)cols.size();
      cols.push_back(c);
    }
    b.colInd.insert(b.colInd.end(), cols.begin(), cols.end());
    b.rowPtr[c + 1] = (int)b.colInd.size();
  }
}

}  // namespace

extern "C" {

const char *phase_hip_mesh_last_error(const phase_hip_mesh *m) { return m ? m->err.c_str() : "null mesh"; }

int phase_hip_mesh_create(phase_hip_mesh **out, int device, int nNodes, const double *xy, int nCells,
                          const int *cptr, const int *cind) {
  if (!out) return PHASE_HIP_EINVAL;
  *out = nullptr;
  phase_hip_mesh *m = new phase_hip_mesh();
  m->device = device;
  int rc = guardedMesh(m, [&] {
    if (nNodes <= 0 || nCells <= 0 || !xy || !cptr || !cind) fail(PHASE_HIP_EINVAL, "empty mesh");
    PH_CHECK(hipSetDevice(device));
    PH_CHECK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    m->nNodes = nNodes;
    m->nCells = nCells;
    // faces in first-encounter order (FiniteVolumeGrid2D::createCell)
    std::unordered_map<unsigned long long, int> dir;
    dir.reserve((size_t)cptr[nCells]);
    std::vector<int> cnt;
    for (int c = 0; c < nCells; ++c) {
      const int b = cptr[c], nv = cptr[c + 1] - b;
      if (nv < 3) fail(PHASE_HIP_EINVAL, "cell " + std::to_string(c) + " has fewer than 3 nodes");
      for (int k = 0; k < nv; ++k) {
        const int n1 = cind[b + k], n2 = cind[b + (k + 1) % nv];
        if (n1 < 0 || n1 >= nNodes || n2 < 0 || n2 >= nNodes) fail(PHASE_HIP_EINVAL, "node index out of range");
        const unsigned long long key = n1 < n2 ? ((unsigned long long)n1 << 32) | (unsigned)n2
                                               : ((unsigned long long)n2 << 32) | (unsigned)n1;
        auto it = dir.find(key);
        if (it == dir.end()) {
          dir.emplace(key, (int)m->fNode1.size());
          m->fNode1.push_back(n1);
          m->fNode2.push_back(n2);
          m->fL.push_back(c);
          m->fR.push_back(-1);
          cnt.push_back(1);
        } else {
          const int f = it->second;
          if (cnt[f] >= 2)
            fail(PHASE_HIP_EINVAL, "an interior face cannot be shared between more than two cells");
          m->fR[f] = c;
          ++cnt[f];
        }
      }
    }
    m->nFaces = (int)m->fNode1.size();
    std::vector<int> ni(nCells, 0), nbd(nCells, 0);
    for (int f = 0; f < m->nFaces; ++f) {
      if (m->fR[f] >= 0) {
        ++ni[m->fL[f]];
        ++ni[m->fR[f]];
        ++m->nInterior;
      } else {
        ++nbd[m->fL[f]];
        ++m->nBoundary;
      }
    }
    m->iptr.assign(nCells + 1, 0);
    m->bptr.assign(nCells + 1, 0);
    for (int c = 0; c < nCells; ++c) {
      m->iptr[c + 1] = m->iptr[c] + ni[c];
      m->bptr[c + 1] = m->bptr[c] + nbd[c];
    }
    m->iface.resize(m->iptr[nCells]);
    m->inb.resize(m->iptr[nCells]);
    m->bface.resize(m->bptr[nCells]);
    std::vector<int> pi(m->iptr.begin(), m->iptr.end() - 1), pb(m->bptr.begin(), m->bptr.end() - 1);
    for (int f = 0; f < m->nFaces; ++f) {
      const int l = m->fL[f], r = m->fR[f];
      if (r >= 0) {
        m->iface[pi[l]] = f;
        m->inb[pi[l]++] = r;
        m->iface[pi[r]] = f;
        m->inb[pi[r]++] = l;
      } else {
        m->bface[pb[l]++] = f;
      }
    }
    hipStream_t st = m->stream;
    m->dXY.upload(std::vector<double>(xy, xy + 2 * (size_t)nNodes), st);
    m->dCptr.upload(std::vector<int>(cptr, cptr + nCells + 1), st);
    m->dCind.upload(std::vector<int>(cind, cind + cptr[nCells]), st);
    m->dFNode1.upload(m->fNode1, st);
    m->dFNode2.upload(m->fNode2, st);
    m->dFL.upload(m->fL, st);
    m->dFR.upload(m->fR, st);
    for (int c = 0; c < nCells; ++c) m->maxDeg = std::max(m->maxDeg, m->iptr[c + 1] - m->iptr[c]);
    m->dIptr.upload(m->iptr, st);
    m->dIface.upload(m->iface.empty() ? std::vector<int>{0} : m->iface, st);
    m->dInb.upload(m->inb.empty() ? std::vector<int>{0} : m->inb, st);
    m->dBptr.upload(m->bptr, st);
    m->dBface.upload(m->bface.empty() ? std::vector<int>{0} : m->bface, st);
    m->dVol.alloc(nCells);
    m->dCC.alloc(2 * (size_t)nCells);
    m->dFC.alloc(2 * (size_t)m->nFaces);
    m->dSf.alloc(2 * (size_t)m->nFaces);
    m->dIDot.alloc(std::max<size_t>(1, m->iface.size()));
    m->dIMs.alloc(std::max<size_t>(1, m->iface.size()));
    m->dBDot.alloc(std::max<size_t>(1, m->bface.size()));
    m->dBMs.alloc(std::max<size_t>(1, m->bface.size()));
    hipLaunchKernelGGL(kCellGeom, dim3(grid1(nCells)), dim3(256), 0, st, nCells, m->dCptr.p, m->dCind.p, m->dXY.p,
                       m->dVol.p, m->dCC.p);
    hipLaunchKernelGGL(kFaceGeom, dim3(grid1(m->nFaces)), dim3(256), 0, st, m->nFaces, m->dFNode1.p, m->dFNode2.p,
                       m->dXY.p, m->dFC.p, m->dSf.p);
    hipLaunchKernelGGL(kLinkGeom, dim3(grid1(nCells)), dim3(256), 0, st, nCells, m->dCC.p, m->dFC.p, m->dSf.p,
                       m->dIptr.p, m->dIface.p, m->dInb.p, m->dBptr.p, m->dBface.p, m->dIDot.p, m->dIMs.p,
                       m->dBDot.p, m->dBMs.p);
    PH_CHECK(hipGetLastError());
    PH_CHECK(hipStreamSynchronize(st));
  });
  if (rc != PHASE_HIP_OK) {
    phase_hip_mesh_destroy(m);
    return rc;
  }
  *out = m;
  return PHASE_HIP_OK;
}

int phase_hip_mesh_destroy(phase_hip_mesh *m) {
  if (!m) return PHASE_HIP_OK;
  (void)hipSetDevice(m->device);
  hipStream_t st = m->stream;
  if (st) (void)hipStreamSynchronize(st);
  delete m;
  if (st) (void)hipStreamDestroy(st);
  return PHASE_HIP_OK;
}

int phase_hip_mesh_sizes(const phase_hip_mesh *m, int *nCells, int *nFaces, int *nInteriorFaces,
                         int *nBoundaryFaces) {
  if (!m) return PHASE_HIP_EINVAL;
  if (nCells) *nCells = m->nCells;
  if (nFaces) *nFaces = m->nFaces;
  if (nInteriorFaces) *nInteriorFaces = m->nInterior;
  if (nBoundaryFaces) *nBoundaryFaces = m->nBoundary;
  return PHASE_HIP_OK;
}

int phase_hip_mesh_faces(const phase_hip_mesh *m, int *lNode, int *rNode, int *lCell, int *rCell) {
  if (!m) return PHASE_HIP_EINVAL;
  if (lNode) std::copy(m->fNode1.begin(), m->fNode1.end(), lNode);
  if (rNode) std::copy(m->fNode2.begin(), m->fNode2.end(), rNode);
  if (lCell) std::copy(m->fL.begin(), m->fL.end(), lCell);
  if (rCell) std::copy(m->fR.begin(), m->fR.end(), rCell);
  return PHASE_HIP_OK;
}

int phase_hip_mesh_geometry(const phase_hip_mesh *mc, double *cellVolume, double *cellCentroid,
                            double *faceCentroid, double *faceSf) {
  phase_hip_mesh *m = const_cast<phase_hip_mesh *>(mc);
  return guardedMesh(m, [&] {
    PH_CHECK(hipSetDevice(m->device));
    hipStream_t st = m->stream;
    if (cellVolume) PH_CHECK(hipMemcpyAsync(cellVolume, m->dVol.p, m->nCells * sizeof(double), hipMemcpyDeviceToHost, st));
    if (cellCentroid)
      PH_CHECK(hipMemcpyAsync(cellCentroid, m->dCC.p, 2 * m->nCells * sizeof(double), hipMemcpyDeviceToHost, st));
    if (faceCentroid)
      PH_CHECK(hipMemcpyAsync(faceCentroid, m->dFC.p, 2 * m->nFaces * sizeof(double), hipMemcpyDeviceToHost, st));
    if (faceSf) PH_CHECK(hipMemcpyAsync(faceSf, m->dSf.p, 2 * m->nFaces * sizeof(double), hipMemcpyDeviceToHost, st));
    PH_CHECK(hipStreamSynchronize(st));
  });
}

}  // extern "C"

namespace {

void bindMesh(phase_hip_solver *s, phase_hip_mesh *m) {
  if (s->boundMesh == m && s->binding && s->havePattern) return;
  solverFreeBinding(s);
  auto *b = new MeshBinding();
  buildBinding(m, *b);
  s->rank = 0;
  std::vector<double> xy;
  if (s->geometric) {
    xy.resize(2 * (size_t)m->nCells);
    PH_CHECK(hipMemcpy(xy.data(), m->dCC.p, xy.size() * sizeof(double), hipMemcpyDeviceToHost));
  }
  solverSetPattern(s, m->nCells, b->rowPtr.data(), b->colInd.data(), xy.empty() ? nullptr : xy.data());
  b->dIslot.upload(b->islot.empty() ? std::vector<int>{0} : b->islot, s->stream);
  b->dIfirst.upload(b->ifirst.empty() ? std::vector<uint8_t>{0} : b->ifirst, s->stream);
  b->dDslot.upload(b->dslot, s->stream);
  b->dRhs.alloc(m->nCells);
  b->dErr.alloc(1);
  b->dErr.zero(s->stream);
  s->binding = b;
  s->boundMesh = m;
  s->haveSolution = false;
}

// Shared body of the host- and device-pointer assembly entry points.  With
// onDevice the operator inputs are device pointers and the boundary-type check
// happens in the kernel (reported by the next solve or export).
void assembleLaplacian(phase_hip_solver *s, phase_hip_mesh *m, int variant, double gamma, const double *gammaFace,
                       const double *gamma0Face, double theta, const int *bcType, const double *phiFace,
                       const double *phi0Cell, const double *phi0Face, bool onDevice) {
  if (!m) fail(PHASE_HIP_EINVAL, "null mesh");
  if (variant < 0 || variant > 3) fail(PHASE_HIP_EINVAL, "laplacian variant must be 0..3");
  if (m->device != s->device) fail(PHASE_HIP_EINVAL, "mesh and solver live on different devices");
  const bool field = variant >= 2, hasTheta = (variant == 0 || variant == 2);
  if (field && !gammaFace) fail(PHASE_HIP_EINVAL, "field-gamma laplacian needs gammaFace");
  if (hasTheta && !phi0Cell) fail(PHASE_HIP_EINVAL, "theta laplacian needs phi0Cell");
  PH_CHECK(hipSetDevice(s->device));
  const int N = m->nCells, F = m->nFaces;
  if (!onDevice && bcType) {
    bool needFixed = false, bad = false;
    for (int f = 0; f < F; ++f) {
      if (m->fR[f] >= 0) continue;
      const int bt = bcType[f];
      if (bt == PHASE_HIP_BC_FIXED) needFixed = true;
      else if (bt != PHASE_HIP_BC_NORMAL_GRADIENT && bt != PHASE_HIP_BC_SYMMETRY) bad = true;
    }
    if (bad) fail(PHASE_HIP_EBOUNDARY, "fv::laplacian: unrecognized or unspecified boundary type.");
    if (needFixed && !phiFace) fail(PHASE_HIP_EINVAL, "FIXED boundaries need phiFace");
    if (needFixed && hasTheta && !phi0Face)
      fail(PHASE_HIP_EINVAL, "theta laplacian with FIXED boundaries needs phi0Face");
  }
  bindMesh(s, m);
  MeshBinding &b = *s->binding;
  hipStream_t st = s->stream;
  PH_CHECK(hipStreamSynchronize(m->stream));
  const double *dG = gammaFace, *dG0 = gamma0Face ? gamma0Face : gammaFace, *dPhiF = phiFace,
               *dPhi0C = phi0Cell, *dPhi0F = phi0Face;
  const int *dBc = bcType;
  if (!onDevice) {
    uploadOpt(b.dStage[0], gammaFace, F, st);
    uploadOpt(b.dStage[1], dG0, F, st);
    uploadOpt(b.dStage[2], phiFace, F, st);
    uploadOpt(b.dStage[3], phi0Cell, N, st);
    uploadOpt(b.dStage[4], phi0Face, F, st);
    dG = gammaFace ? b.dStage[0].p : nullptr;
    dG0 = dG0 ? b.dStage[1].p : nullptr;
    dPhiF = phiFace ? b.dStage[2].p : nullptr;
    dPhi0C = phi0Cell ? b.dStage[3].p : nullptr;
    dPhi0F = phi0Face ? b.dStage[4].p : nullptr;
  }
  if (bcType) {
    b.dBc.alloc(F);
    PH_CHECK(hipMemcpyAsync(b.dBc.p, bcType, F * sizeof(int),
                            onDevice ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    dBc = b.dBc.p;
  } else {
    b.dBc.release();
    dBc = nullptr;
  }
  b.bcOnHost = false;
  hipLaunchKernelGGL(kAssembleLaplacian, dim3(grid1(N)), dim3(256), 0, st, N, variant, gamma, field ? dG : nullptr,
                     (field && hasTheta) ? dG0 : nullptr, theta, dBc, dPhiF, hasTheta ? dPhi0C : nullptr,
                     hasTheta ? dPhi0F : nullptr, m->dIptr.p, m->dIface.p, m->dInb.p, m->dIDot.p, m->dIMs.p,
                     b.dIslot.p, b.dIfirst.p, m->dBptr.p, m->dBface.p, m->dBDot.p, m->dBMs.p, b.dDslot.p, s->dVals.p,
                     b.dRhs.p, b.dErr.p);
  PH_CHECK(hipGetLastError());
  solverValuesChanged(s);
  hipLaunchKernelGGL(kRhsToB, dim3(grid1(N)), dim3(256), 0, st, N, s->dPerm.p, b.dRhs.p, s->dB.p);
  PH_CHECK(hipGetLastError());
  s->haveRhs = true;
  s->lastVariant = variant;
  if (!onDevice) checkAssembly(s);
}

void subRhs(phase_hip_solver *s, const double *v, bool onDevice) {
  if (!s->binding) fail(PHASE_HIP_ESTATE, "no assembled equation");
  PH_CHECK(hipSetDevice(s->device));
  const int N = s->rank;
  MeshBinding &b = *s->binding;
  const double *dv = v;
  if (!onDevice) {
    b.dStage[5].alloc(N);
    PH_CHECK(hipMemcpyAsync(b.dStage[5].p, v, N * sizeof(double), hipMemcpyHostToDevice, s->stream));
    dv = b.dStage[5].p;
  }
  hipLaunchKernelGGL(kSubRhs, dim3(grid1(N)), dim3(256), 0, s->stream, N, s->dPerm.p, dv, b.dRhs.p, s->dB.p);
  PH_CHECK(hipGetLastError());
  if (!onDevice) PH_CHECK(hipStreamSynchronize(s->stream));
}

}  // namespace

namespace phase {

// Report a boundary-type error flagged by the assembly kernel (and clear it).
void checkAssembly(phase_hip_solver *s) {
  if (!s->binding) return;
  MeshBinding &b = *s->binding;
  int e = 0;
  PH_CHECK(hipMemcpyAsync(&e, b.dErr.p, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  PH_CHECK(hipStreamSynchronize(s->stream));
  if (e) {
    b.dErr.zero(s->stream);
    fail(PHASE_HIP_EBOUNDARY, "fv::laplacian: unrecognized or unspecified boundary type.");
  }
}

}  // namespace phase

extern "C" {

int phase_hip_assemble_laplacian(phase_hip_solver *s, phase_hip_mesh *m, int variant, double gamma,
                                 const double *gammaFace, const double *gamma0Face, double theta,
                                 const int *bcType, const double *phiFace, const double *phi0Cell,
                                 const double *phi0Face) {
  return guardedSolver(s, [&] {
    assembleLaplacian(s, m, variant, gamma, gammaFace, gamma0Face, theta, bcType, phiFace, phi0Cell, phi0Face,
                      false);
  });
}

int phase_hip_assemble_laplacian_device(phase_hip_solver *s, phase_hip_mesh *m, int variant, double gamma,
                                        const double *gammaFace, const double *gamma0Face, double theta,
                                        const int *bcType, const double *phiFace, const double *phi0Cell,
                                        const double *phi0Face) {
  return guardedSolver(s, [&] {
    assembleLaplacian(s, m, variant, gamma, gammaFace, gamma0Face, theta, bcType, phiFace, phi0Cell, phi0Face,
                      true);
  });
}

int phase_hip_equation_sub_rhs(phase_hip_solver *s, const double *v) {
  return guardedSolver(s, [&] { subRhs(s, v, false); });
}

int phase_hip_equation_sub_rhs_device(phase_hip_solver *s, const double *v) {
  return guardedSolver(s, [&] { subRhs(s, v, true); });
}

int phase_hip_equation_export_padded(phase_hip_solver *s, int *nslots, int *rowPtr, int *colInd, double *vals,
                                     double *rhs) {
  return guardedSolver(s, [&] {
    if (!s->binding || !s->boundMesh) fail(PHASE_HIP_ESTATE, "no assembled equation");
    PH_CHECK(hipSetDevice(s->device));
    phase_hip_mesh *m = s->boundMesh;
    MeshBinding &b = *s->binding;
    const int N = m->nCells;
    const bool pFirst = s->lastVariant >= 2;
    checkAssembly(s);
    if (!b.bcOnHost) {
      b.lastBc.clear();
      if (b.dBc.p) {
        b.lastBc.resize(m->nFaces);
        PH_CHECK(hipMemcpyAsync(b.lastBc.data(), b.dBc.p, m->nFaces * sizeof(int), hipMemcpyDeviceToHost, s->stream));
        PH_CHECK(hipStreamSynchronize(s->stream));
      }
      b.bcOnHost = true;
    }
    // per row: first-touch column order of the last variant (CrsEquation::addCoeff)
    std::vector<int> rp(N + 1, 0);
    std::vector<std::vector<int>> order(N);
    for (int c = 0; c < N; ++c) {
      auto &cols = order[c];
      for (int k = m->iptr[c]; k < m->iptr[c + 1]; ++k) {
        const int nb = m->inb[k];
        auto add = [&](int col) {
          if (std::find(cols.begin(), cols.end(), col) == cols.end()) cols.push_back(col);
        };
        if (pFirst) { add(c); add(nb); }
        else { add(nb); add(c); }
      }
      bool fixed = false;
      if (!b.lastBc.empty())
        for (int k = m->bptr[c]; k < m->bptr[c + 1]; ++k)
          if (b.lastBc[m->bface[k]] == PHASE_HIP_BC_FIXED) fixed = true;
      if (fixed && std::find(cols.begin(), cols.end(), c) == cols.end()) cols.push_back(c);
      rp[c + 1] = rp[c] + std::max<int>(5, (int)cols.size());
    }
    if (nslots) *nslots = rp[N];
    if (!rowPtr) return;
    std::vector<double> hv(s->dVals.n), hr(N);
    PH_CHECK(hipMemcpyAsync(hv.data(), s->dVals.p, hv.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    PH_CHECK(hipMemcpyAsync(hr.data(), b.dRhs.p, N * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    PH_CHECK(hipStreamSynchronize(s->stream));
    std::copy(rp.begin(), rp.end(), rowPtr);
    for (int c = 0; c < N; ++c) {
      int w = rp[c];
      for (int col : order[c]) {
        int slot = -1;
        for (int j = b.rowPtr[c]; j < b.rowPtr[c + 1]; ++j)
          if (b.colInd[j] == col) slot = j;
        colInd[w] = col;
        vals[w++] = hv[slot];
      }
      for (; w < rp[c + 1]; ++w) {
        colInd[w] = -1;
        vals[w] = 0.;
      }
    }
    if (rhs) std::copy(hr.begin(), hr.end(), rhs);
  });
}

}  // extern "C"
