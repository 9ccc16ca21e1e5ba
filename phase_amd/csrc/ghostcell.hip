// Ghost-cell immersed boundary rows (src/2D/Unstructured/FiniteVolume/ImmersedBoundary/
// GhostCellImmersedBoundary.cpp:12-90, GhostCellImmersedBoundaryBcStencil.cpp:3-70,
// GhostCellImmersedBoundaryFixedBcStencil.cpp:17-22,
// GhostCellImmersedBoundaryNormalGradientBcStencil.cpp:22-42) for circle and box bodies.
//
// Set-up, host, once per mesh and object set (GhostCellImmersedBoundary::updateCells):
//   per object in order, the remaining domain cells whose centroid lies
//   strictly inside its shape are solid; those with a face or node neighbour
//   outside it (internalPerimeterCells with diagonals) are ghost cells;
//   ghost-cell stencil: bp = nearestIntersect(x_c), ip = 2 bp - x_c,
//   nw = nearestEdgeUnitNormal(bp) (the Body queries below), the four cells of the
//   mesh node nearest to ip (lowest node id on a tie) and Ainv, the inverse of
//   their bilinear rows (x y, x, y, 1) (Gauss-Jordan with partial pivoting;
//   the reference calls LAPACK getrf/getri).
// Device, per call: a fresh equation whose ghost rows hold the FIXED or
// NORMAL_GRADIENT stencil, whose solid rows hold a unit diagonal, and whose
// fluid rows are empty; callers remove the non-fluid rows of their fv::
// equations (remove(cell)) and add these.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "equation.h"
#include "mesh.h"

using namespace phase;

// per ghost cell, for the force: the cell, its body, bp, ip, nw, the image
// point's interpolation cells and Ainv, and the boundary point's fit cells
// with the inverse of their bilinear matrix
struct GhostPoint {
  int cell, body;
  double bp[2], ip[2], nw[2];
  int ipCells[4], bpCells[4];
  double ipInv[16], bpInv[16];
};

struct phase_hip_gcib {
  phase_hip_mesh *mesh = nullptr;
  int N = 0, nObj = 0;
  std::vector<int> types, owner;
  std::vector<int> kind;      // per body: PHASE_HIP_SHAPE_CIRCLE or _BOX
  std::vector<double> geom;   // 4 per body: cx, cy, r, 0 or lower x, lower y, upper x, upper y
  std::vector<GhostPoint> points;
  std::vector<double> cc;     // cell centroids (2 per cell), host copy taken at creation
  DBuf<int> dGather;     // 9 cells per ghost point: cell, ipCells, bpCells
  DBuf<double> dValues;  // p, u.x, u.y at those cells
  DBuf<int> dTypes, dStencil, dFixCols, dNgCols;  // per cell: stencil index or -1; 5 columns per stencil
  DBuf<double> dFixVals, dNgVals;
  std::string err;
};

namespace {

inline int g256g(long n) { return (int)std::max<long>(1, std::min<long>((n + 255) / 256, 8192)); }

template <class T>
std::vector<T> fetch(const DBuf<T> &d, size_t n, hipStream_t st) {
  std::vector<T> h(n);
  if (n) PH_CHECK(hipMemcpyAsync(h.data(), d.p, n * sizeof(T), hipMemcpyDeviceToHost, st));
  PH_CHECK(hipStreamSynchronize(st));
  return h;
}

// inverse of a 4x4 row-major matrix, Gauss-Jordan with partial pivoting
bool inverse4(const double *a, double *inv) {
  double m[4][8];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      m[i][j] = a[4 * i + j];
      m[i][4 + j] = i == j ? 1. : 0.;
    }
  for (int c = 0; c < 4; ++c) {
    int p = c;
    for (int i = c + 1; i < 4; ++i)
      if (std::fabs(m[i][c]) > std::fabs(m[p][c])) p = i;
    if (!(std::fabs(m[p][c]) > 0.)) return false;
    if (p != c)
      for (int j = 0; j < 8; ++j) std::swap(m[p][j], m[c][j]);
    const double d = m[c][c];
    for (int j = 0; j < 8; ++j) m[c][j] /= d;
    for (int i = 0; i < 4; ++i)
      if (i != c) {
        const double f = m[i][c];
        if (f != 0.)
          for (int j = 0; j < 8; ++j) m[i][j] -= f * m[c][j];
      }
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) inv[4 * i + j] = m[i][4 + j];
  return true;
}

// Shape queries of the immersed bodies: Circle.cpp (isInside, nearestIntersect),
// Box.cpp (isInside, nearestIntersect, nearestEdge, edges) and
// ImmersedBoundaryObject::nearestEdgeUnitNormal (ImmersedBoundaryObject.cpp:76-90).
struct Body {
  int kind;
  double g[4];
  bool inside(double x, double y) const {
    if (kind == PHASE_HIP_SHAPE_CIRCLE) {
      const double dx = x - g[0], dy = y - g[1];
      return dx * dx + dy * dy < g[2] * g[2];
    }
    return x < g[2] && x > g[0] && y < g[3] && y > g[1];
  }
  void centre(double &x, double &y) const {
    if (kind == PHASE_HIP_SHAPE_CIRCLE) {
      x = g[0];
      y = g[1];
    } else {
      x = (g[0] + g[2]) / 2.;
      y = (g[1] + g[3]) / 2.;
    }
  }
  void edge(int k, double *a, double *b) const {  // Box::edges order
    const double px[4] = {g[0], g[2], g[2], g[0]}, py[4] = {g[1], g[1], g[3], g[3]};
    a[0] = px[k];
    a[1] = py[k];
    b[0] = px[(k + 1) % 4];
    b[1] = py[(k + 1) % 4];
  }
  void nearestIntersect(double px, double py, double &bx, double &by) const {
    if (kind == PHASE_HIP_SHAPE_CIRCLE) {
      const double dx = px - g[0], dy = py - g[1], dm = std::sqrt(dx * dx + dy * dy);
      bx = dx / dm * g[2] + g[0];
      by = dy / dm * g[2] + g[1];
      return;
    }
    double best = INFINITY;
    for (int k = 0; k < 4; ++k) {
      double a[2], b[2];
      edge(k, a, b);
      const double tx = b[0] - a[0], ty = b[1] - a[1], rx = px - a[0], ry = py - a[1];
      const double d = rx * tx + ry * ty, ms = tx * tx + ty * ty;
      double qx, qy;
      if (d >= 0. && d <= ms) {  // LineSegment2D::isBounded
        qx = (a[0] + (d * tx) / ms) - px;
        qy = (a[1] + (d * ty) / ms) - py;
      } else {
        const double ax = a[0] - px, ay = a[1] - py, cx = b[0] - px, cy = b[1] - py;
        const bool useA = ax * ax + ay * ay < cx * cx + cy * cy;
        qx = (useA ? a[0] : b[0]) - px;
        qy = (useA ? a[1] : b[1]) - py;
      }
      const double d2 = qx * qx + qy * qy;
      if (d2 < best) {
        best = d2;
        bx = px + qx;
        by = py + qy;
      }
    }
  }
  void wallNormal(double bx, double by, double &nx, double &ny) const {
    if (kind == PHASE_HIP_SHAPE_CIRCLE) {
      const double wx = g[0] - bx, wy = g[1] - by, wm = std::sqrt(wx * wx + wy * wy);
      nx = wx / wm;
      ny = wy / wm;
      return;
    }
    double best = INFINITY, ea[2] = {0., 0.}, eb[2] = {0., 0.};
    for (int k = 0; k < 4; ++k) {
      double a[2], b[2];
      edge(k, a, b);
      const double tx = b[0] - a[0], ty = b[1] - a[1], rx = bx - a[0], ry = by - a[1];
      const double d = rx * tx + ry * ty, ms = tx * tx + ty * ty;
      double d2;
      if (d >= 0. && d <= ms) {
        const double qx = rx - (d * tx) / ms, qy = ry - (d * ty) / ms;
        d2 = qx * qx + qy * qy;
      } else {
        const double ax = a[0] - bx, ay = a[1] - by, cx = b[0] - bx, cy = b[1] - by;
        d2 = std::min(ax * ax + ay * ay, cx * cx + cy * cy);
      }
      if (d2 < best) {
        best = d2;
        ea[0] = a[0];
        ea[1] = a[1];
        eb[0] = b[0];
        eb[1] = b[1];
      }
    }
    const double tx = eb[0] - ea[0], ty = eb[1] - ea[1];
    const double n0 = ty, n1 = -tx;  // LineSegment2D::norm = (B - A).normalVec()
    double cx, cy;
    centre(cx, cy);
    const double mx = cx - (ea[0] + eb[0]) / 2., my = cy - (ea[1] + eb[1]) / 2.;
    const double m = std::sqrt(n0 * n0 + n1 * n1);
    const double ux = n0 / m, uy = n1 / m;
    if (n0 * mx + n1 * my > 0.) {
      nx = ux;
      ny = uy;
    } else {
      nx = -ux;
      ny = -uy;
    }
  }
};

// p, u.x, u.y at the gathered cells
__global__ void kGatherPU(int m, const int *__restrict__ idx, const double *__restrict__ p,
                          const double *__restrict__ u, double *__restrict__ out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    const int c = idx[i];
    out[3 * i] = p[c];
    out[3 * i + 1] = u[2 * c];
    out[3 * i + 2] = u[2 * c + 1];
  }
}

// one fresh row per (component, cell): stencil, unit diagonal or empty
__global__ void kEqGhostRows(int n, int ncomp, int stride, const int *__restrict__ types,
                             const int *__restrict__ stencil, const int *__restrict__ scols,
                             const double *__restrict__ svals, const double *__restrict__ src, int *cols,
                             double *vals, int *cnt, double *rhs) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    for (int q = 0; q < ncomp; ++q) {
      const long row = (long)q * n + c;
      const long base = row * stride;
      for (int j = 0; j < stride; ++j) {
        cols[base + j] = -1;
        vals[base + j] = 0.;
      }
      int used = 0;
      auto add = [&](int col, double v) {  // CrsEquation::addCoeff, first touch
        for (int j = 0; j < used; ++j)
          if (cols[base + j] == col) {
            vals[base + j] += v;
            return;
          }
        cols[base + used] = col;
        vals[base + used] = v;
        ++used;
      };
      const int t = types[c];
      if (t == 1) {
        const int s = stencil[c];
        for (int j = 0; j < 5; ++j) {
          const int k = scols[5 * s + j];
          if (k >= 0) add(q * n + k, svals[5 * s + j]);
        }
      } else if (t == 2) {
        add((int)row, 1.);
      }
      cnt[row] = used;
      rhs[row] = t != 0 ? 0. + src[(long)c * ncomp + q] : 0.;
    }
  }
}

}  // namespace

extern "C" {

int phase_hip_gcib_destroy(phase_hip_gcib *ib) {
  delete ib;
  return PHASE_HIP_OK;
}

const char *phase_hip_gcib_last_error(const phase_hip_gcib *ib) { return ib ? ib->err.c_str() : "null object"; }

int phase_hip_gcib_create_shapes(phase_hip_gcib **out, phase_hip_mesh *m, int nBodies, const int *kinds,
                                 const double *geometry) {
  if (!out || !m || nBodies < 0 || (nBodies > 0 && (!kinds || !geometry))) return PHASE_HIP_EINVAL;
  const int nCircles = nBodies;
  *out = nullptr;
  auto *ib = new phase_hip_gcib();
  try {
    PH_CHECK(hipSetDevice(m->device));
    hipStream_t st = m->stream;
    const int N = m->nCells;
    ib->mesh = m;
    ib->N = N;
    ib->nObj = nCircles;
    ib->cc = fetch(m->dCC, 2 * (size_t)N, st);
    const std::vector<double> &cc = ib->cc;
    const std::vector<double> xy = fetch(m->dXY, 2 * (size_t)m->nNodes, st);
    const std::vector<int> cptr = fetch(m->dCptr, N + 1, st);
    const std::vector<int> cind = fetch(m->dCind, cptr[N], st);
    std::vector<std::vector<int>> nodeCells(m->nNodes);
    for (int c = 0; c < N; ++c)
      for (int k = cptr[c]; k < cptr[c + 1]; ++k) nodeCells[cind[k]].push_back(c);
    // Cell::cellLinks: face neighbours and diagonal (node) neighbours
    auto forLinks = [&](int c, auto &&fn) {
      for (int e = m->iptr[c]; e < m->iptr[c + 1]; ++e) fn(m->inb[e]);
      for (int k = cptr[c]; k < cptr[c + 1]; ++k)
        for (int kc : nodeCells[cind[k]])
          if (kc != c) fn(kc);
    };
    ib->types.assign(N, 0);
    ib->owner.assign(N, -1);
    std::vector<char> inside(N);
    std::vector<Body> bodies(nBodies);
    for (int o = 0; o < nBodies; ++o) {
      Body &bd = bodies[o];
      bd.kind = kinds[o];
      std::copy(geometry + 4 * o, geometry + 4 * o + 4, bd.g);
      if (bd.kind == PHASE_HIP_SHAPE_CIRCLE) {
        if (!(bd.g[2] > 0.)) fail(PHASE_HIP_EINVAL, "circle radius must be positive");
      } else if (bd.kind == PHASE_HIP_SHAPE_BOX) {
        // Box::Box orders the corners
        if (bd.g[2] < bd.g[0]) std::swap(bd.g[0], bd.g[2]);
        if (bd.g[3] < bd.g[1]) std::swap(bd.g[1], bd.g[3]);
        if (!(bd.g[2] > bd.g[0]) || !(bd.g[3] > bd.g[1])) fail(PHASE_HIP_EINVAL, "empty box");
      } else {
        fail(PHASE_HIP_EINVAL, "body shapes are PHASE_HIP_SHAPE_CIRCLE or PHASE_HIP_SHAPE_BOX");
      }
      ib->kind.push_back(bd.kind);
      ib->geom.insert(ib->geom.end(), bd.g, bd.g + 4);
    }
    for (int o = 0; o < nBodies; ++o) {
      const Body &bd = bodies[o];
      for (int c = 0; c < N; ++c) inside[c] = bd.inside(cc[2 * c], cc[2 * c + 1]);
      for (int c = 0; c < N; ++c) {
        if (!inside[c] || ib->types[c] != 0) continue;
        bool ghost = false;
        forLinks(c, [&](int k) { ghost |= !inside[k]; });
        ib->types[c] = ghost ? 1 : 2;
        ib->owner[c] = o;
      }
    }
    // nearest node: bins over the nodes
    double x0 = xy[0], x1 = xy[0], y0 = xy[1], y1 = xy[1];
    for (int i = 0; i < m->nNodes; ++i) {
      x0 = std::min(x0, xy[2 * i]);
      x1 = std::max(x1, xy[2 * i]);
      y0 = std::min(y0, xy[2 * i + 1]);
      y1 = std::max(y1, xy[2 * i + 1]);
    }
    const long nb = std::max<long>(1, std::min<long>(2048, (long)std::sqrt((double)m->nNodes)));
    const double bw = (x1 - x0) / nb + 1e-300, bh = (y1 - y0) / nb + 1e-300;
    auto bx = [&](double v) { return std::min<long>(nb - 1, std::max<long>(0, (long)((v - x0) / bw))); };
    auto by = [&](double v) { return std::min<long>(nb - 1, std::max<long>(0, (long)((v - y0) / bh))); };
    std::vector<int> bptr(nb * nb + 1, 0), bnode(m->nNodes);
    for (int i = 0; i < m->nNodes; ++i) ++bptr[by(xy[2 * i + 1]) * nb + bx(xy[2 * i]) + 1];
    for (long b = 0; b < nb * nb; ++b) bptr[b + 1] += bptr[b];
    {
      std::vector<int> fill(bptr.begin(), bptr.end() - 1);
      for (int i = 0; i < m->nNodes; ++i) bnode[fill[by(xy[2 * i + 1]) * nb + bx(xy[2 * i])]++] = i;
    }
    auto nearestNode = [&](double px, double py) {
      const long ix = bx(px), iy = by(py);
      int best = -1;
      double bd = 0.;
      for (long ring = 0; ring <= nb; ++ring) {
        for (long j = iy - ring; j <= iy + ring; ++j)
          for (long i = ix - ring; i <= ix + ring; ++i) {
            if (j < 0 || i < 0 || j >= nb || i >= nb) continue;
            if (std::max(std::labs(j - iy), std::labs(i - ix)) != ring) continue;
            for (int q = bptr[j * nb + i]; q < bptr[j * nb + i + 1]; ++q) {
              const int k = bnode[q];
              const double dx = xy[2 * k] - px, dy = xy[2 * k + 1] - py;
              const double d = dx * dx + dy * dy;
              if (best < 0 || d < bd || (d == bd && k < best)) {
                best = k;
                bd = d;
              }
            }
          }
        // every node outside the searched rings is at least ring * bin away
        if (best >= 0 && (double)ring * std::min(bw, bh) > std::sqrt(bd)) break;
      }
      return best;
    };
    std::vector<int> stencil(N, -1), fixCols, ngCols;
    std::vector<double> fixVals, ngVals;
    for (int c = 0; c < N; ++c) {
      if (ib->types[c] != 1) continue;
      const int o = ib->owner[c];
      const double xc = cc[2 * c], yc = cc[2 * c + 1];
      double bpx = 0., bpy = 0., nwx = 0., nwy = 0.;
      bodies[o].nearestIntersect(xc, yc, bpx, bpy);
      const double ipx = 2. * bpx - xc, ipy = 2. * bpy - yc;
      bodies[o].wallNormal(bpx, bpy, nwx, nwy);
      const int node = nearestNode(ipx, ipy);
      const std::vector<int> &cells = nodeCells[node];
      if (cells.size() != 4)
        fail(PHASE_HIP_EINVAL, "ghost-cell stencil of cell " + std::to_string(c) + ": nearest node has " +
                                   std::to_string(cells.size()) + " cells (four needed)");
      double A[16], Ai[16];
      for (int i = 0; i < 4; ++i) {
        const double x = cc[2 * cells[i]], y = cc[2 * cells[i] + 1];
        A[4 * i] = x * y;
        A[4 * i + 1] = x;
        A[4 * i + 2] = y;
        A[4 * i + 3] = 1.;
      }
      if (!inverse4(A, Ai)) fail(PHASE_HIP_EINVAL, "singular ghost-cell interpolation at cell " + std::to_string(c));
      const double v[4] = {ipx * ipy, ipx, ipy, 1.};
      double va[4];
      for (int j = 0; j < 4; ++j) {
        double s = 0.;
        for (int k = 0; k < 4; ++k) s += v[k] * Ai[4 * k + j];
        va[j] = s;
      }
      {  // the force's data (applyHydrodynamicForce)
        GhostPoint g{};
        g.cell = c;
        g.body = o;
        g.bp[0] = bpx;
        g.bp[1] = bpy;
        g.ip[0] = ipx;
        g.ip[1] = ipy;
        g.nw[0] = nwx;
        g.nw[1] = nwy;
        std::copy(cells.begin(), cells.end(), g.ipCells);
        std::copy(Ai, Ai + 16, g.ipInv);
        const std::vector<int> &bcells = nodeCells[nearestNode(bpx, bpy)];
        if (bcells.size() != 4)
          fail(PHASE_HIP_EINVAL, "boundary-point fit of cell " + std::to_string(c) + " needs a node with four cells");
        double B[16];
        for (int i = 0; i < 4; ++i) {
          const double x = cc[2 * bcells[i]], y = cc[2 * bcells[i] + 1];
          B[4 * i] = x * y;
          B[4 * i + 1] = x;
          B[4 * i + 2] = y;
          B[4 * i + 3] = 1.;
          g.bpCells[i] = bcells[i];
        }
        if (!inverse4(B, g.bpInv)) fail(PHASE_HIP_EINVAL, "singular boundary-point fit at cell " + std::to_string(c));
        ib->points.push_back(g);
      }
      stencil[c] = (int)(fixCols.size() / 5);
      for (int j = 0; j < 4; ++j) {
        fixCols.push_back(cells[j]);
        fixVals.push_back(va[j] / 2.);
      }
      fixCols.push_back(c);
      fixVals.push_back(0.5);
      if (std::find(cells.begin(), cells.end(), c) != cells.end()) {
        const double g[4] = {bpy * nwx + bpx * nwy, nwx, nwy, 0.};
        for (int j = 0; j < 4; ++j) {
          double s = 0.;
          for (int k = 0; k < 4; ++k) s += g[k] * Ai[4 * k + j];
          ngCols.push_back(cells[j]);
          ngVals.push_back(s);
        }
        ngCols.push_back(-1);
        ngVals.push_back(0.);
      } else {
        const double ex = ipx - xc, ey = ipy - yc, l = std::sqrt(ex * ex + ey * ey);
        for (int j = 0; j < 4; ++j) {
          ngCols.push_back(cells[j]);
          ngVals.push_back(-va[j] / l);
        }
        ngCols.push_back(c);
        ngVals.push_back(1. / l);
      }
    }
    std::vector<int> gather;
    for (const GhostPoint &g : ib->points) {
      gather.push_back(g.cell);
      gather.insert(gather.end(), g.ipCells, g.ipCells + 4);
      gather.insert(gather.end(), g.bpCells, g.bpCells + 4);
    }
    ib->dGather.upload(gather, st);
    ib->dValues.alloc(3 * gather.size());
    ib->dTypes.upload(ib->types, st);
    ib->dStencil.upload(stencil, st);
    ib->dFixCols.upload(fixCols, st);
    ib->dFixVals.upload(fixVals, st);
    ib->dNgCols.upload(ngCols, st);
    ib->dNgVals.upload(ngVals, st);
    PH_CHECK(hipStreamSynchronize(st));
  } catch (const HipError &x) {
    delete ib;
    return x.code;
  } catch (const std::exception &) {
    delete ib;
    return PHASE_HIP_EINVAL;
  }
  *out = ib;
  return PHASE_HIP_OK;
}

int phase_hip_gcib_create(phase_hip_gcib **out, phase_hip_mesh *m, int nCircles, const double *circles) {
  if (nCircles < 0 || (nCircles > 0 && !circles)) return PHASE_HIP_EINVAL;
  std::vector<int> kinds(nCircles, PHASE_HIP_SHAPE_CIRCLE);
  std::vector<double> geom(4 * (size_t)nCircles, 0.);
  for (int o = 0; o < nCircles; ++o) std::copy(circles + 3 * o, circles + 3 * o + 3, geom.begin() + 4 * o);
  return phase_hip_gcib_create_shapes(out, m, nCircles, kinds.data(), geom.data());
}

int phase_hip_gcib_hydrodynamic_force(phase_hip_gcib *ib, double rho, double mu, const double *uCell,
                                      const double *pCell, double gx, double gy, double *force) {
  if (!ib || !uCell || !pCell || !force) return PHASE_HIP_EINVAL;
  try {
    phase_hip_mesh *m = ib->mesh;
    PH_CHECK(hipSetDevice(m->device));
    const int ng = (int)ib->dGather.n;
    std::vector<double> v(3 * (size_t)ng);
    if (ng) {
      hipLaunchKernelGGL(kGatherPU, dim3(g256g(ng)), dim3(256), 0, m->stream, ng, ib->dGather.p, pCell, uCell,
                         ib->dValues.p);
      PH_CHECK(hipGetLastError());
      PH_CHECK(hipMemcpyAsync(v.data(), ib->dValues.p, v.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    }
    PH_CHECK(hipStreamSynchronize(m->stream));
    struct Stress {
      double angle, x, y, pr, sh;
    };
    for (int o = 0; o < ib->nObj; ++o) {
      Body bd{ib->kind[o], {ib->geom[4 * o], ib->geom[4 * o + 1], ib->geom[4 * o + 2], ib->geom[4 * o + 3]}};
      double cx, cy;
      bd.centre(cx, cy);
      std::vector<Stress> pts;
      for (size_t k = 0; k < ib->points.size(); ++k) {
        const GhostPoint &g = ib->points[k];
        if (g.body != o) continue;
        const double *val = &v[27 * k];  // 9 cells x (p, ux, uy)
        // bpValue(p) = (p_c + ipValue(p)) / 2, ipValue = (ip.x ip.y, ip.x, ip.y, 1) Ainv p_cells
        const double w[4] = {g.ip[0] * g.ip[1], g.ip[0], g.ip[1], 1.};
        double ipv = 0.;
        for (int j = 0; j < 4; ++j) {
          double cj = 0.;
          for (int i = 0; i < 4; ++i) cj += w[i] * g.ipInv[4 * i + j];
          ipv += cj * val[3 * (1 + j)];
        }
        const double pr = (val[0] + ipv) / 2. + rho * (g.bp[0] * gx + g.bp[1] * gy);
        // bpGrad(u): X = A^-1 u(bpCells); rows d/dx = (bp.y, 1, 0, 0) X, d/dy = (bp.x, 0, 1, 0) X
        double X[4][2];
        for (int i = 0; i < 4; ++i)
          for (int q = 0; q < 2; ++q) {
            double t = 0.;
            for (int j = 0; j < 4; ++j) t += g.bpInv[4 * i + j] * val[3 * (5 + j) + 1 + q];
            X[i][q] = t;
          }
        double G[2][2];
        for (int q = 0; q < 2; ++q) {
          G[0][q] = g.bp[1] * X[0][q] + X[1][q];
          G[1][q] = g.bp[0] * X[0][q] + X[2][q];
        }
        // Tensor2D(xx, xy, yx, yy) = (G00, G01, G10, G11); dot(T, nw); dot(., nw.tangentVec())
        const double tnx = G[0][0] * g.nw[0] + G[0][1] * g.nw[1];
        const double tny = G[1][0] * g.nw[0] + G[1][1] * g.nw[1];
        const double sh = mu * (tnx * -g.nw[1] + tny * g.nw[0]);
        pts.push_back({std::atan2(g.bp[1] - cy, g.bp[0] - cx), g.bp[0], g.bp[1], pr, sh});
      }
      std::sort(pts.begin(), pts.end(), [](const Stress &a, const Stress &b) { return a.angle < b.angle; });
      double fx = 0., fy = 0.;
      for (size_t i = 0; i < pts.size(); ++i) {
        const Stress &a = pts[i], &b = pts[(i + 1) % pts.size()];
        const double dx = b.x - a.x, dy = b.y - a.y;
        const double pm = -(a.pr + b.pr) / 2., sm = (a.sh + b.sh) / 2.;
        fx += pm * dy + sm * dx;  // normalVec() = (dy, -dx)
        fy += pm * -dx + sm * dy;
      }
      force[2 * o] = fx;
      force[2 * o + 1] = fy;
    }
  } catch (const HipError &x) {
    ib->err = x.msg;
    return x.code;
  } catch (const std::exception &x) {
    ib->err = x.what();
    return PHASE_HIP_EINVAL;
  }
  return PHASE_HIP_OK;
}

int phase_hip_gcib_cell_types(const phase_hip_gcib *ib, int *types, int *owner, int onDevice) {
  if (!ib) return PHASE_HIP_EINVAL;
  try {
    hipStream_t st = ib->mesh->stream;
    if (types) {
      if (onDevice)
        PH_CHECK(hipMemcpyAsync(types, ib->dTypes.p, ib->N * sizeof(int), hipMemcpyDeviceToDevice, st));
      else
        std::copy(ib->types.begin(), ib->types.end(), types);
    }
    if (owner) std::copy(ib->owner.begin(), ib->owner.end(), owner);
    PH_CHECK(hipStreamSynchronize(st));
  } catch (const HipError &x) {
    return x.code;
  }
  return PHASE_HIP_OK;
}

}  // extern "C"

namespace {

int ghostRows(phase_hip_equation *e, const phase_hip_gcib *ib, int ncomp, bool fixed, const std::vector<double> &src) {
  if (!e || !ib) return PHASE_HIP_EINVAL;
  try {
    if (e->mesh != ib->mesh) fail(PHASE_HIP_EINVAL, "equation and immersed boundary of different meshes");
    if (e->ncomp != ncomp) fail(PHASE_HIP_EINVAL, ncomp == 1 ? "bcs needs a scalar equation" :
                                                              "velocityBcs needs a Vector2D equation");
    if (e->stride < 5) fail(PHASE_HIP_EINVAL, "equation rows hold fewer than five entries");
    phase_hip_mesh *m = e->mesh;
    PH_CHECK(hipSetDevice(m->device));
    DBuf<double> dSrc;
    dSrc.upload(src, m->stream);
    hipLaunchKernelGGL(kEqGhostRows, dim3(g256g(m->nCells)), dim3(256), 0, m->stream, m->nCells, ncomp, e->stride,
                       ib->dTypes.p, ib->dStencil.p, fixed ? ib->dFixCols.p : ib->dNgCols.p,
                       fixed ? ib->dFixVals.p : ib->dNgVals.p, dSrc.p, e->cols.p, e->vals.p, e->cnt.p, e->rhs.p);
    PH_CHECK(hipGetLastError());
    PH_CHECK(hipStreamSynchronize(m->stream));  // before dSrc is freed
    e->compact = false;
  } catch (const HipError &x) {
    e->err = x.msg;
    return x.code;
  } catch (const std::exception &x) {
    e->err = x.what();
    return PHASE_HIP_EINVAL;
  }
  return PHASE_HIP_OK;
}

}  // namespace

extern "C" {

int phase_hip_eq_gcib_bcs(phase_hip_equation *e, const phase_hip_gcib *ib, int bcType, double refValue) {
  if (!ib || !e) return PHASE_HIP_EINVAL;
  if (bcType != PHASE_HIP_BC_FIXED && bcType != PHASE_HIP_BC_NORMAL_GRADIENT) {
    e->err = "GhostCellImmersedBoundary::bcs: invalid boundary condition type";
    return PHASE_HIP_EBOUNDARY;
  }
  const bool fixed = bcType == PHASE_HIP_BC_FIXED;
  std::vector<double> src(ib->N, 0.);
  for (int c = 0; c < ib->N; ++c)
    if (ib->types[c] == 1 || (ib->types[c] == 2 && fixed)) src[c] = -refValue;
  return ghostRows(e, ib, 1, fixed, src);
}

int phase_hip_eq_gcib_motion_bcs(phase_hip_equation *e, const phase_hip_gcib *ib, const double *objectMotion) {
  if (!ib || !e || (ib->nObj > 0 && !objectMotion)) return PHASE_HIP_EINVAL;
  // Motion::velocity(pt) = vel + omega (pt - pos).tangentVec() (Motion.h:31-33),
  // pos the body's centroid: at bp for the ghost rows, the centroid for the solid rows
  std::vector<double> bpOf(2 * (size_t)ib->N, 0.);
  for (const GhostPoint &g : ib->points) {
    bpOf[2 * g.cell] = g.bp[0];
    bpOf[2 * g.cell + 1] = g.bp[1];
  }
  const std::vector<double> &cc = ib->cc;
  std::vector<double> src(2 * (size_t)ib->N, 0.);
  for (int c = 0; c < ib->N; ++c)
    if (ib->types[c] != 0) {
      const int o = ib->owner[c];
      Body bd{ib->kind[o], {ib->geom[4 * o], ib->geom[4 * o + 1], ib->geom[4 * o + 2], ib->geom[4 * o + 3]}};
      double px, py;
      bd.centre(px, py);
      const double x = ib->types[c] == 1 ? bpOf[2 * c] : cc[2 * c];
      const double y = ib->types[c] == 1 ? bpOf[2 * c + 1] : cc[2 * c + 1];
      const double w = objectMotion[3 * o + 2];
      const double ux = objectMotion[3 * o] + -(y - py) * w, uy = objectMotion[3 * o + 1] + (x - px) * w;
      src[2 * c] = -ux;
      src[2 * c + 1] = -uy;
    }
  return ghostRows(e, ib, 2, true, src);
}

int phase_hip_eq_gcib_velocity_bcs(phase_hip_equation *e, const phase_hip_gcib *ib, const double *objectVelocity) {
  if (!ib || !e || (ib->nObj > 0 && !objectVelocity)) return PHASE_HIP_EINVAL;
  std::vector<double> motion(3 * (size_t)ib->nObj, 0.);
  for (int o = 0; o < ib->nObj; ++o) {
    motion[3 * o] = objectVelocity[2 * o];
    motion[3 * o + 1] = objectVelocity[2 * o + 1];
  }
  return phase_hip_eq_gcib_motion_bcs(e, ib, motion.data());
}

}  // extern "C"
