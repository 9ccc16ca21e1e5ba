// TEST INFRASTRUCTURE ONLY -- FP64 CPU oracle of the Phase discretise-and-solve
// path.  Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg as the checker / CPU baseline; never linked into the product library.
//
// Restated from the reference sources (paths relative to the reference root).
// Compile with -ffp-contract=off: every expression below is evaluated in the
// same operation order as the reference line it cites, so results are the
// same IEEE binary64 roundings (the reference binary itself is built with
// GCC's default contraction, CMakeLists.txt:10, so bit-identity with *that*
// binary is not claimed).
//
// Parity status (how each part is pinned; see tests/):
//   mesh topology/links ......... closed form on rectilinear meshes + invariants
//   geometry .................... closed form (dyadic meshes) + shoelace/centroid
//                                 invariants; Boost.Geometry bits unpinned
//   CrsEquation semantics ....... hand-worked known-answer tests
//   fv:: assembly ............... closed form on dyadic meshes, dense brute force,
//                                 symmetry/row-sum invariants
//   SpMV ........................ dense matvec (numpy) to rounding
//   ILU(0) ...................... product L*U reproduces A on the pattern
//   BiCGStab+ILU(0)/ILUT ........ true residual <= tol and scipy splu solution
#include "phase_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace {

struct V2 {
  double x, y;
};
inline V2 operator-(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
inline V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
// Vector2D.cpp:145-147
inline double dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
// Vector2D.h:18
inline double magSqr(V2 a) { return a.x * a.x + a.y * a.y; }
inline double mag(V2 a) { return std::sqrt(a.x * a.x + a.y * a.y); }

struct Mesh {
  int nNodes = 0, nCells = 0, nFaces = 0, nInterior = 0, nBoundary = 0;
  std::vector<V2> node;
  std::vector<int> cptr, cind;
  std::vector<int> fNode1, fNode2, fL, fR;  // face nodes, lCell, rCell (-1)
  std::vector<V2> fC, fSf;                  // centroid, unnormalised normal
  std::vector<double> vol;
  std::vector<V2> cC;
  // links in the reference's order: interior links ascending face id, boundary
  // links ascending face id (FiniteVolumeGrid2D.cpp:396-418).
  std::vector<int> iptr, iface, inb, bptr, bface;
};

// Face::outwardNorm(point), Face.cpp:48-50
inline V2 outwardNorm(const Mesh &m, int f, V2 p) {
  V2 n = m.fSf[f];
  return dot(m.fC[f] - p, n) > 0. ? n : V2{-n.x, -n.y};
}

struct Eqn {
  std::vector<int> rowPtr, colInd;
  std::vector<double> vals, rhs;
  Eqn(int nRows, int nnz)
      : rowPtr(nRows + 1), colInd((size_t)nRows * nnz, -1),
        vals((size_t)nRows * nnz, 0.), rhs(nRows, 0.) {
    // CrsEquation.cpp:9-14
    for (int i = 0; i <= nRows; ++i) rowPtr[i] = i * nnz;
  }
  int rank() const { return (int)rowPtr.size() - 1; }

  // CrsEquation::addCoeff / setCoeff, CrsEquation.cpp:113-151: first touch of
  // an empty (-1) slot claims it; a full row grows by one slot at its end.
  void touch(int row, int col, double val, bool add) {
    for (int j = rowPtr[row]; j < rowPtr[row + 1]; ++j) {
      if (colInd[j] == col) {
        if (add) vals[j] += val; else vals[j] = val;
        return;
      } else if (colInd[j] == -1) {
        colInd[j] = col;
        vals[j] = val;
        return;
      }
    }
    colInd.insert(colInd.begin() + rowPtr[row + 1], col);
    vals.insert(vals.begin() + rowPtr[row + 1], val);
    for (size_t r = row + 1; r < rowPtr.size(); ++r) ++rowPtr[r];
  }
  void addCoeff(int row, int col, double val) { touch(row, col, val, true); }

  // CrsEquation::operator+= / -=, CrsEquation.cpp:185-275: rows are rebuilt,
  // zero values and empty slots are dropped, rhs entries accumulate into a
  // matching column or are appended.
  void merge(const Eqn &o, int sign) {
    std::vector<int> rp(1, 0), ci;
    std::vector<double> v;
    for (int row = 0; row < rank(); ++row) {
      rp.push_back(rp.back());
      for (int j = rowPtr[row]; j < rowPtr[row + 1]; ++j) {
        if (vals[j] == 0. || colInd[j] < 0) continue;
        ci.push_back(colInd[j]);
        v.push_back(vals[j]);
        ++rp.back();
      }
      for (int j = o.rowPtr[row]; j < o.rowPtr[row + 1]; ++j) {
        if (o.vals[j] == 0. || o.colInd[j] < 0) continue;
        int k = rp[row];
        for (; k < rp[row + 1]; ++k)
          if (ci[k] == o.colInd[j]) break;
        if (k != rp[row + 1]) {
          if (sign > 0) v[k] += o.vals[j]; else v[k] -= o.vals[j];
        } else {
          ci.push_back(o.colInd[j]);
          v.push_back(sign > 0 ? o.vals[j] : -o.vals[j]);
          ++rp.back();
        }
      }
    }
    rowPtr = rp;
    colInd = ci;
    vals = v;
    for (int i = 0; i < rank(); ++i) {
      if (sign > 0) rhs[i] += o.rhs[i]; else rhs[i] -= o.rhs[i];
    }
  }
};

void setErr(char *err, int len, const std::string &msg) {
  if (err && len > 0) {
    std::snprintf(err, len, "%s", msg.c_str());
  }
}

// Row index of (cell, component): IndexMap block layout, IndexMap.cpp:31-36.
inline int rowOf(const Mesh &m, int cell, int comp) { return comp * m.nCells + cell; }

// FiniteVolumeEquation<T>::add(cell, nb, Scalar): the same coefficient goes to
// every component block (ScalarFiniteVolumeEquation.cpp:17-21,
// VectorFiniteVolumeEquation.cpp:21-29).
inline void addS(Eqn &e, const Mesh &m, int ncomp, int cell, int nb, double v) {
  for (int c = 0; c < ncomp; ++c) e.addCoeff(rowOf(m, cell, c), rowOf(m, nb, c), v);
}
// addSource(cell, T): component-wise (VectorFiniteVolumeEquation.cpp:104-109).
inline void addSrc(Eqn &e, const Mesh &m, int ncomp, int cell, const double *v) {
  for (int c = 0; c < ncomp; ++c) e.rhs[rowOf(m, cell, c)] += v[c];
}
// add(cell, nb, Tensor2D), VectorFiniteVolumeEquation.cpp:48-66.
inline void addT(Eqn &e, const Mesh &m, int cell, int nb, double xx, double xy,
                 double yx, double yy) {
  e.addCoeff(rowOf(m, cell, 0), rowOf(m, nb, 0), xx);
  if (xy != 0.) e.addCoeff(rowOf(m, cell, 0), rowOf(m, nb, 1), xy);
  if (yx != 0.) e.addCoeff(rowOf(m, cell, 1), rowOf(m, nb, 0), yx);
  e.addCoeff(rowOf(m, cell, 1), rowOf(m, nb, 1), yy);
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- mesh
// FiniteVolumeGrid2D::init(nodes,cptr,cind) (FiniteVolumeGrid2D.cpp:20-35),
// createCell (:84-114), init() (:396-418); Cell ctor (Cell.cpp:8-29); Face
// ctor (Face.cpp:9-18); links (CellLink.cpp:8, InteriorLink.cpp:9,
// BoundaryLink.cpp:8).
void *orc_mesh_create(int nNodes, const double *xy, int nCells, const int *cptr,
                      const int *cind, char *err, int errLen) {
  Mesh *m = new Mesh();
  m->nNodes = nNodes;
  m->nCells = nCells;
  m->node.resize(nNodes);
  for (int i = 0; i < nNodes; ++i) m->node[i] = {xy[2 * i], xy[2 * i + 1]};
  m->cptr.assign(cptr, cptr + nCells + 1);
  m->cind.assign(cind, cind + cptr[nCells]);
  m->vol.resize(nCells);
  m->cC.resize(nCells);

  std::map<std::pair<int, int>, int> dir;
  std::vector<int> cnt;
  for (int c = 0; c < nCells; ++c) {
    int b = cptr[c], nv = cptr[c + 1] - cptr[c];
    // Polygon area (shoelace) and centroid (Bashein-Detmer): the formulas
    // behind boost::geometry::area / centroid called at Polygon.cpp:218-231.
    double a2 = 0., sx = 0., sy = 0.;
    for (int k = 0; k < nv; ++k) {
      V2 p = m->node[cind[b + k]], q = m->node[cind[b + (k + 1) % nv]];
      double ai = p.x * q.y - q.x * p.y;
      a2 += ai;
      sx += (p.x + q.x) * ai;
      sy += (p.y + q.y) * ai;
    }
    if (a2 < 0.) {  // boost::geometry::correct reverses a clockwise ring
      a2 = -a2; sx = -sx; sy = -sy;
    }
    m->vol[c] = a2 / 2.;
    m->cC[c] = {sx / (3. * a2), sy / (3. * a2)};
    for (int k = 0; k < nv; ++k) {
      int n1 = cind[b + k], n2 = cind[b + (k + 1) % nv];
      auto key = n1 < n2 ? std::make_pair(n1, n2) : std::make_pair(n2, n1);
      auto it = dir.find(key);
      if (it == dir.end()) {
        int f = m->nFaces++;
        dir[key] = f;
        m->fNode1.push_back(n1);
        m->fNode2.push_back(n2);
        m->fL.push_back(c);
        m->fR.push_back(-1);
        cnt.push_back(1);
        V2 a = m->node[n1], bb = m->node[n2];
        V2 s = a + bb;
        m->fC.push_back({0.5 * s.x, 0.5 * s.y});
        V2 t = bb - a;
        m->fSf.push_back({t.y, -t.x});  // Vector2D::normalVec, Vector2D.h:26
      } else {
        int f = it->second;
        if (cnt[f] >= 2) {
          setErr(err, errLen, "an interior face cannot be shared between more than two cells");
          delete m;
          return nullptr;
        }
        m->fR[f] = c;
        ++cnt[f];
      }
    }
  }
  // Links: interior faces add a link to both cells, boundary faces to lCell,
  // all in ascending face id.
  std::vector<int> ni(nCells, 0), nb(nCells, 0);
  for (int f = 0; f < m->nFaces; ++f) {
    if (m->fR[f] >= 0) { ++ni[m->fL[f]]; ++ni[m->fR[f]]; ++m->nInterior; }
    else { ++nb[m->fL[f]]; ++m->nBoundary; }
  }
  m->iptr.assign(nCells + 1, 0);
  m->bptr.assign(nCells + 1, 0);
  for (int c = 0; c < nCells; ++c) {
    m->iptr[c + 1] = m->iptr[c] + ni[c];
    m->bptr[c + 1] = m->bptr[c] + nb[c];
  }
  m->iface.resize(m->iptr[nCells]);
  m->inb.resize(m->iptr[nCells]);
  m->bface.resize(m->bptr[nCells]);
  std::vector<int> pi(m->iptr.begin(), m->iptr.end() - 1), pb(m->bptr.begin(), m->bptr.end() - 1);
  for (int f = 0; f < m->nFaces; ++f) {
    int l = m->fL[f], r = m->fR[f];
    if (r >= 0) {
      m->iface[pi[l]] = f; m->inb[pi[l]++] = r;
      m->iface[pi[r]] = f; m->inb[pi[r]++] = l;
    } else {
      m->bface[pb[l]++] = f;
    }
  }
  return m;
}

void orc_mesh_free(void *mesh) { delete static_cast<Mesh *>(mesh); }

void orc_mesh_sizes(const void *mesh, int *nFaces, int *nInteriorFaces, int *nBoundaryFaces) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  *nFaces = m.nFaces;
  *nInteriorFaces = m.nInterior;
  *nBoundaryFaces = m.nBoundary;
}

void orc_mesh_faces(const void *mesh, int *lNode, int *rNode, int *lCell, int *rCell,
                    double *centroid, double *sf) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  for (int f = 0; f < m.nFaces; ++f) {
    lNode[f] = m.fNode1[f]; rNode[f] = m.fNode2[f];
    lCell[f] = m.fL[f]; rCell[f] = m.fR[f];
    centroid[2 * f] = m.fC[f].x; centroid[2 * f + 1] = m.fC[f].y;
    sf[2 * f] = m.fSf[f].x; sf[2 * f + 1] = m.fSf[f].y;
  }
}

void orc_mesh_cells(const void *mesh, double *volume, double *centroid) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  for (int c = 0; c < m.nCells; ++c) {
    volume[c] = m.vol[c];
    centroid[2 * c] = m.cC[c].x; centroid[2 * c + 1] = m.cC[c].y;
  }
}

void orc_mesh_links(const void *mesh, int *iptr, int *iface, int *inb, int *bptr, int *bface) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  std::copy(m.iptr.begin(), m.iptr.end(), iptr);
  std::copy(m.iface.begin(), m.iface.end(), iface);
  std::copy(m.inb.begin(), m.inb.end(), inb);
  std::copy(m.bptr.begin(), m.bptr.end(), bptr);
  std::copy(m.bface.begin(), m.bface.end(), bface);
}

// ---------------------------------------------------------------- equations
void *orc_eqn_create(int nRows, int nnzPerRow) { return new Eqn(nRows, nnzPerRow); }
void *orc_eqn_copy(const void *eqn) { return new Eqn(*static_cast<const Eqn *>(eqn)); }
void orc_eqn_free(void *eqn) { delete static_cast<Eqn *>(eqn); }
int orc_eqn_rank(const void *eqn) { return static_cast<const Eqn *>(eqn)->rank(); }
int orc_eqn_nslots(const void *eqn) { return (int)static_cast<const Eqn *>(eqn)->colInd.size(); }
void orc_eqn_get(const void *eqn, int *rowPtr, int *colInd, double *vals, double *rhs) {
  const Eqn &e = *static_cast<const Eqn *>(eqn);
  std::copy(e.rowPtr.begin(), e.rowPtr.end(), rowPtr);
  std::copy(e.colInd.begin(), e.colInd.end(), colInd);
  std::copy(e.vals.begin(), e.vals.end(), vals);
  std::copy(e.rhs.begin(), e.rhs.end(), rhs);
}
void orc_eqn_add_coeff(void *eqn, int row, int col, double val) {
  static_cast<Eqn *>(eqn)->touch(row, col, val, true);
}
void orc_eqn_set_coeff(void *eqn, int row, int col, double val) {
  static_cast<Eqn *>(eqn)->touch(row, col, val, false);
}
void orc_eqn_add_rhs(void *eqn, int row, double val) { static_cast<Eqn *>(eqn)->rhs[row] += val; }
void orc_eqn_plus_eqn(void *lhs, const void *rhs, int sign) {
  static_cast<Eqn *>(lhs)->merge(*static_cast<const Eqn *>(rhs), sign);
}
// CrsEquation::operator-=(Vector) / operator==(Vector), CrsEquation.cpp:282-285,308-311
void orc_eqn_minus_vector(void *eqn, const double *v) {
  Eqn &e = *static_cast<Eqn *>(eqn);
  for (int i = 0; i < e.rank(); ++i) e.rhs[i] -= v[i];
}
// CrsEquation::operator==(Scalar), CrsEquation.cpp:299-303
void orc_eqn_eq_scalar(void *eqn, double s) {
  Eqn &e = *static_cast<Eqn *>(eqn);
  if (s != 0.)
    for (int i = 0; i < e.rank(); ++i) e.rhs[i] -= s;
}
// CrsEquation::operator*=(Scalar), CrsEquation.cpp:287-291
void orc_eqn_scale(void *eqn, double s) {
  Eqn &e = *static_cast<Eqn *>(eqn);
  for (double &v : e.vals) v *= s;
  for (double &v : e.rhs) v *= s;
}
// operator*(ScalarFiniteVolumeField, FiniteVolumeEquation<T>)
// (FiniteVolumeEquation.h:100-102, VectorFiniteVolumeEquation.cpp:164-170):
// scale(cell, field(cell)) = CrsEquation::scaleRow (CrsEquation.cpp:153-159)
// of every row of the cell, values and rhs; rows are comp * nCells + cell.
void orc_eqn_scale_cells(void *eqn, int nCells, const double *field) {
  Eqn &e = *static_cast<Eqn *>(eqn);
  for (int row = 0; row < e.rank(); ++row) {
    const double f = field[row % nCells];
    for (int j = e.rowPtr[row]; j < e.rowPtr[row + 1]; ++j) e.vals[j] *= f;
    e.rhs[row] *= f;
  }
}
// remove(cell) (ScalarFiniteVolumeEquation.cpp:38-43,
// VectorFiniteVolumeEquation.cpp:78-91): the cell's rows lose their columns
// (-1, values kept) and their rhs
void orc_eqn_remove_cells(void *eqn, int nCells, const int *mask) {
  Eqn &e = *static_cast<Eqn *>(eqn);
  for (int row = 0; row < e.rank(); ++row) {
    if (!mask[row % nCells]) continue;
    for (int j = e.rowPtr[row]; j < e.rowPtr[row + 1]; ++j) e.colInd[j] = -1;
    e.rhs[row] = 0.;
  }
}
// Implicit under-relaxation as preserved in the comment of
// ScalarFiniteVolumeEquation.cpp:45-55 / VectorFiniteVolumeEquation.cpp:129-146:
//   a_P /= alpha;  rhs -= (1 - alpha) * a_P * phi_old   (new a_P)
// phi holds one value per row (block layout for vector equations).
void orc_eqn_relax(void *eqn, double alpha, const double *phi) {
  Eqn &e = *static_cast<Eqn *>(eqn);
  for (int row = 0; row < e.rank(); ++row) {
    for (int j = e.rowPtr[row]; j < e.rowPtr[row + 1]; ++j) {
      if (e.colInd[j] == row) {
        e.vals[j] /= alpha;
        e.rhs[row] -= (1. - alpha) * e.vals[j] * phi[row];
        break;
      }
    }
  }
}

// ---------------------------------------------------------------- operators
void *orc_laplacian(const void *mesh, int variant, int ncomp, double gamma,
                    const double *gammaFace, const double *gamma0Face, double theta,
                    const int *bcType, const double *phiFace, const double *phi0Cell,
                    const double *phi0Face, char *err, int errLen) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  const bool hasTheta = (variant == 0 || variant == 2);
  const bool fieldGamma = (variant == 2 || variant == 3);
  if (ncomp == 2 && !hasTheta) {
    // the non-theta Vector2D forms use the generic template: identical to the
    // scalar case per component, SYMMETRY contributing nothing.
  }
  Eqn *e = new Eqn(ncomp * m.nCells, 5);
  for (int c = 0; c < m.nCells; ++c) {
    V2 cc = m.cC[c];
    for (int k = m.iptr[c]; k < m.iptr[c + 1]; ++k) {
      int f = m.iface[k], nb = m.inb[k];
      V2 rc = m.cC[nb] - cc;
      V2 sf = outwardNorm(m, f, cc);
      double g = fieldGamma ? gammaFace[f] : gamma;
      double coeff = g * dot(rc, sf) / magSqr(rc);
      if (!hasTheta) {
        if (!fieldGamma) {  // Laplacian.h:57-58
          addS(*e, m, ncomp, c, nb, coeff);
          addS(*e, m, ncomp, c, c, -coeff);
        } else {            // Laplacian.h:141-142
          addS(*e, m, ncomp, c, c, -coeff);
          addS(*e, m, ncomp, c, nb, coeff);
        }
        continue;
      }
      double coeff0 = coeff;
      if (fieldGamma) coeff0 = gamma0Face[f] * dot(rc, sf) / magSqr(rc);
      if (!fieldGamma) {    // Laplacian.h:17-21
        addS(*e, m, ncomp, c, nb, theta * coeff);
        addS(*e, m, ncomp, c, c, theta * -coeff);
      } else {              // Laplacian.h:99-101 / Laplacian.cpp:75-77
        addS(*e, m, ncomp, c, c, theta * -coeff);
        addS(*e, m, ncomp, c, nb, theta * coeff);
      }
      double s[2];
      for (int q = 0; q < ncomp; ++q)
        s[q] = (1. - theta) * coeff0 * (phi0Cell[ncomp * nb + q] - phi0Cell[ncomp * c + q]);
      addSrc(*e, m, ncomp, c, s);
    }
    for (int k = m.bptr[c]; k < m.bptr[c + 1]; ++k) {
      int f = m.bface[k];
      V2 rf = m.fC[f] - cc;
      V2 sf = outwardNorm(m, f, cc);
      double g = fieldGamma ? gammaFace[f] : gamma;
      double coeff = g * dot(rf, sf) / magSqr(rf);
      double coeff0 = coeff;
      if (fieldGamma && hasTheta) coeff0 = gamma0Face[f] * dot(rf, sf) / magSqr(rf);
      int bt = bcType ? bcType[f] : ORC_BC_NORMAL_GRADIENT;
      if (bt == ORC_BC_FIXED) {
        double s[2];
        if (!hasTheta) {    // Laplacian.h:69-72
          addS(*e, m, ncomp, c, c, -coeff);
          for (int q = 0; q < ncomp; ++q) s[q] = coeff * phiFace[ncomp * f + q];
          addSrc(*e, m, ncomp, c, s);
        } else {            // Laplacian.h:31-36
          addS(*e, m, ncomp, c, c, theta * -coeff);
          for (int q = 0; q < ncomp; ++q) s[q] = theta * coeff * phiFace[ncomp * f + q];
          addSrc(*e, m, ncomp, c, s);
          for (int q = 0; q < ncomp; ++q)
            s[q] = (1. - theta) * coeff0 * (phi0Face[ncomp * f + q] - phi0Cell[ncomp * c + q]);
          addSrc(*e, m, ncomp, c, s);
        }
      } else if (bt == ORC_BC_NORMAL_GRADIENT) {
      } else if (bt == ORC_BC_SYMMETRY) {
        if (ncomp == 2 && hasTheta) {  // Laplacian.cpp:36-43
          V2 tv = {-sf.y, sf.x};       // tangentVec, Vector2D.h:28
          double tm = mag(tv);
          V2 tw = {tv.x / tm, tv.y / tm};
          addS(*e, m, ncomp, c, c, theta * -coeff);
          double tc = theta * coeff;
          addT(*e, m, c, c, tc * (tw.x * tw.x), tc * (tw.x * tw.y), tc * (tw.y * tw.x),
               tc * (tw.y * tw.y));
          double p0x = phi0Cell[2 * c], p0y = phi0Cell[2 * c + 1];
          double d = p0x * tw.x + p0y * tw.y;
          double w = (1. - theta) * coeff0;
          double s[2] = {w * (tw.x * d - p0x), w * (tw.y * d - p0y)};
          addSrc(*e, m, ncomp, c, s);
        }
      } else {
        setErr(err, errLen, "fv::laplacian: unrecognized or unspecified boundary type.");
        delete e;
        return nullptr;
      }
    }
  }
  return e;
}

// fv::ddt(field, timeStep), TimeDerivative.h:37-48
void *orc_ddt(const void *mesh, int ncomp, double timeStep, const double *phi0Cell) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  Eqn *e = new Eqn(ncomp * m.nCells, 5);
  for (int c = 0; c < m.nCells; ++c) {
    addS(*e, m, ncomp, c, c, m.vol[c] / timeStep);
    double s[2];
    for (int q = 0; q < ncomp; ++q) s[q] = -m.vol[c] * phi0Cell[ncomp * c + q] / timeStep;
    addSrc(*e, m, ncomp, c, s);
  }
  return e;
}

// fv::ddt(Scalar rho, field, dt) (TimeDerivative.h:7-19; rhoCell == null) and
// fv::ddt(rho field, field, dt) (:21-35: rho for the coefficient, rho0 for
// the source); cellMask (optional) restricts the rows to a cell group as
// ddt(field, dt, cells) does (:50-62)
void *orc_ddt_rho(const void *mesh, int ncomp, double timeStep, double rho, const double *rhoCell,
                  const double *rho0Cell, const double *phi0Cell, const int *cellMask) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  Eqn *e = new Eqn(ncomp * m.nCells, 5);
  for (int c = 0; c < m.nCells; ++c) {
    if (cellMask && !cellMask[c]) continue;
    double s[2];
    if (rhoCell) {
      addS(*e, m, ncomp, c, c, rhoCell[c] * m.vol[c] / timeStep);
      for (int q = 0; q < ncomp; ++q) s[q] = -rho0Cell[c] * m.vol[c] * phi0Cell[ncomp * c + q] / timeStep;
    } else {
      addS(*e, m, ncomp, c, c, rho * m.vol[c] / timeStep);
      for (int q = 0; q < ncomp; ++q) s[q] = -rho * m.vol[c] * phi0Cell[ncomp * c + q] / timeStep;
    }
    addSrc(*e, m, ncomp, c, s);
  }
  return e;
}

// fv::div(u, phi, theta) upwind, Divergence.h:7-53
void *orc_div(const void *mesh, int ncomp, const double *uFace, const double *u0Face,
              double theta, const int *bcType, const double *phiFace, const double *phi0Cell,
              const double *phi0Face, char *err, int errLen) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  Eqn *e = new Eqn(ncomp * m.nCells, 5);
  for (int c = 0; c < m.nCells; ++c) {
    V2 cc = m.cC[c];
    for (int k = m.iptr[c]; k < m.iptr[c + 1]; ++k) {
      int f = m.iface[k], nb = m.inb[k];
      V2 sf = outwardNorm(m, f, cc);
      double flux = dot(V2{uFace[2 * f], uFace[2 * f + 1]}, sf);
      double flux0 = dot(V2{u0Face[2 * f], u0Face[2 * f + 1]}, sf);
      addS(*e, m, ncomp, c, c, theta * std::max(flux, 0.));
      addS(*e, m, ncomp, c, nb, theta * std::min(flux, 0.));
      double s[2];
      for (int q = 0; q < ncomp; ++q) s[q] = (1. - theta) * std::max(flux0, 0.) * phi0Cell[ncomp * c + q];
      addSrc(*e, m, ncomp, c, s);
      for (int q = 0; q < ncomp; ++q) s[q] = (1. - theta) * std::min(flux0, 0.) * phi0Cell[ncomp * nb + q];
      addSrc(*e, m, ncomp, c, s);
    }
    for (int k = m.bptr[c]; k < m.bptr[c + 1]; ++k) {
      int f = m.bface[k];
      V2 sf = outwardNorm(m, f, cc);
      double flux = dot(V2{uFace[2 * f], uFace[2 * f + 1]}, sf);
      double flux0 = dot(V2{u0Face[2 * f], u0Face[2 * f + 1]}, sf);
      int bt = bcType ? bcType[f] : ORC_BC_NORMAL_GRADIENT;
      double s[2];
      if (bt == ORC_BC_FIXED) {
        for (int q = 0; q < ncomp; ++q) s[q] = theta * flux * phiFace[ncomp * f + q];
        addSrc(*e, m, ncomp, c, s);
        for (int q = 0; q < ncomp; ++q) s[q] = (1. - theta) * flux0 * phi0Face[ncomp * f + q];
        addSrc(*e, m, ncomp, c, s);
      } else if (bt == ORC_BC_NORMAL_GRADIENT) {
        addS(*e, m, ncomp, c, c, theta * flux);
        for (int q = 0; q < ncomp; ++q) s[q] = (1. - theta) * flux0 * phi0Cell[ncomp * c + q];
        addSrc(*e, m, ncomp, c, s);
      } else if (bt == ORC_BC_SYMMETRY) {
      } else {
        setErr(err, errLen, "fv::div: unrecognized or unspecified boundary type.");
        delete e;
        return nullptr;
      }
    }
  }
  return e;
}

// fv::divc(u, phi, theta), central, Divergence.h:55-103.  The boundary
// flux0 uses u, not u0, as the reference does.
void *orc_divc(const void *mesh, int ncomp, const double *uFace, const double *u0Face, double theta,
               const int *bcType, const double *phiFace, const double *phi0Cell, const double *phi0Face, char *err,
               int errLen) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  Eqn *e = new Eqn(ncomp * m.nCells, 5);
  for (int c = 0; c < m.nCells; ++c) {
    const V2 cc = m.cC[c];
    double s[2];
    for (int k = m.iptr[c]; k < m.iptr[c + 1]; ++k) {
      const int f = m.iface[k], nb = m.inb[k];
      const V2 sf = outwardNorm(m, f, cc);
      const double flux = theta * dot(V2{uFace[2 * f], uFace[2 * f + 1]}, sf);
      const double flux0 = (1. - theta) * dot(V2{u0Face[2 * f], u0Face[2 * f + 1]}, sf);
      const double lc = mag(m.fC[f] - cc), ln = mag(m.fC[f] - m.cC[nb]);
      const double g = ln / (lc + ln);
      addS(*e, m, ncomp, c, c, g * flux);
      addS(*e, m, ncomp, c, nb, (1. - g) * flux);
      for (int q = 0; q < ncomp; ++q)
        s[q] = flux0 * (g * phi0Cell[ncomp * c + q] + (1. - g) * phi0Cell[ncomp * nb + q]);
      addSrc(*e, m, ncomp, c, s);
    }
    for (int k = m.bptr[c]; k < m.bptr[c + 1]; ++k) {
      const int f = m.bface[k];
      const V2 sf = outwardNorm(m, f, cc);
      const double flux = theta * dot(V2{uFace[2 * f], uFace[2 * f + 1]}, sf);
      const double flux0 = (1. - theta) * dot(V2{uFace[2 * f], uFace[2 * f + 1]}, sf);
      const int bt = bcType ? bcType[f] : ORC_BC_NORMAL_GRADIENT;
      if (bt == ORC_BC_FIXED) {
        for (int q = 0; q < ncomp; ++q) s[q] = flux * phiFace[ncomp * f + q];
        addSrc(*e, m, ncomp, c, s);
        for (int q = 0; q < ncomp; ++q) s[q] = flux0 * phi0Face[ncomp * f + q];
        addSrc(*e, m, ncomp, c, s);
      } else if (bt == ORC_BC_NORMAL_GRADIENT) {
        addS(*e, m, ncomp, c, c, flux);
        for (int q = 0; q < ncomp; ++q) s[q] = flux0 * phi0Cell[ncomp * c + q];
        addSrc(*e, m, ncomp, c, s);
      } else if (bt == ORC_BC_SYMMETRY) {
      } else {
        setErr(err, errLen, "fv::divc: unrecognized or unspecified boundary type.");
        delete e;
        return nullptr;
      }
    }
  }
  return e;
}

// fv::dive(u, phi, theta), explicit, ExplicitDivergence.h:7-51: sources only,
// from the two older levels (u0, phi0) and (u1, phi1)
void *orc_dive(const void *mesh, int ncomp, const double *u0Face, const double *u1Face, double theta,
               const int *bcType, const double *phi0Cell, const double *phi1Cell, const double *phi0Face,
               const double *phi1Face, char *err, int errLen) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  Eqn *e = new Eqn(ncomp * m.nCells, 5);
  for (int c = 0; c < m.nCells; ++c) {
    const V2 cc = m.cC[c];
    double s[2];
    for (int k = m.iptr[c]; k < m.iptr[c + 1]; ++k) {
      const int f = m.iface[k], nb = m.inb[k];
      const V2 sf = outwardNorm(m, f, cc);
      const double flux0 = dot(V2{u0Face[2 * f], u0Face[2 * f + 1]}, sf);
      const double flux1 = dot(V2{u1Face[2 * f], u1Face[2 * f + 1]}, sf);
      for (int q = 0; q < ncomp; ++q) s[q] = theta * std::max(flux0, 0.) * phi0Cell[ncomp * c + q];
      addSrc(*e, m, ncomp, c, s);
      for (int q = 0; q < ncomp; ++q) s[q] = theta * std::min(flux0, 0.) * phi0Cell[ncomp * nb + q];
      addSrc(*e, m, ncomp, c, s);
      for (int q = 0; q < ncomp; ++q) s[q] = (1. - theta) * std::max(flux1, 0.) * phi1Cell[ncomp * c + q];
      addSrc(*e, m, ncomp, c, s);
      for (int q = 0; q < ncomp; ++q) s[q] = (1. - theta) * std::min(flux1, 0.) * phi1Cell[ncomp * nb + q];
      addSrc(*e, m, ncomp, c, s);
    }
    for (int k = m.bptr[c]; k < m.bptr[c + 1]; ++k) {
      const int f = m.bface[k];
      const V2 sf = outwardNorm(m, f, cc);
      const double flux0 = dot(V2{u0Face[2 * f], u0Face[2 * f + 1]}, sf);
      const double flux1 = dot(V2{u1Face[2 * f], u1Face[2 * f + 1]}, sf);
      const int bt = bcType ? bcType[f] : ORC_BC_NORMAL_GRADIENT;
      if (bt == ORC_BC_FIXED || bt == ORC_BC_NORMAL_GRADIENT) {
        for (int q = 0; q < ncomp; ++q) s[q] = theta * flux0 * phi0Face[ncomp * f + q];
        addSrc(*e, m, ncomp, c, s);
        for (int q = 0; q < ncomp; ++q) s[q] = (1. - theta) * flux1 * phi1Face[ncomp * f + q];
        addSrc(*e, m, ncomp, c, s);
      } else if (bt == ORC_BC_SYMMETRY) {
      } else {
        setErr(err, errLen, "fv::dive: unrecognized or unspecified boundary type.");
        delete e;
        return nullptr;
      }
    }
  }
  return e;
}

// fv::div(phiU, gradU, u), Divergence.cpp:3-47: upwind with the upwind cell's
// gradient (a Jacobian per cell, xx xy yx yy; dot(J, r) = (xx r.x + xy r.y,
// yx r.x + yy r.y), Tensor2D.cpp:48-50) carried to the face; Vector2D only
void *orc_div_grad(const void *mesh, const double *phiUFace, const double *gradU, const int *bcType,
                   const double *uFace, char *err, int errLen) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  Eqn *e = new Eqn(2 * m.nCells, 5);
  auto jdot = [&](int c, V2 r, double flux, double *s) {
    const double *J = gradU + 4 * c;
    const double vx = J[0] * r.x + J[1] * r.y, vy = J[2] * r.x + J[3] * r.y;
    s[0] = flux * vx;
    s[1] = flux * vy;
  };
  for (int c = 0; c < m.nCells; ++c) {
    const V2 cc = m.cC[c];
    double s[2];
    for (int k = m.iptr[c]; k < m.iptr[c + 1]; ++k) {
      const int f = m.iface[k], nb = m.inb[k];
      const double flux = dot(V2{phiUFace[2 * f], phiUFace[2 * f + 1]}, outwardNorm(m, f, cc));
      if (flux > 0.) {
        addS(*e, m, 2, c, c, flux);
        jdot(c, m.fC[f] - cc, flux, s);
      } else {
        addS(*e, m, 2, c, nb, flux);
        jdot(nb, m.fC[f] - m.cC[nb], flux, s);
      }
      addSrc(*e, m, 2, c, s);
    }
    for (int k = m.bptr[c]; k < m.bptr[c + 1]; ++k) {
      const int f = m.bface[k];
      const double flux = dot(V2{phiUFace[2 * f], phiUFace[2 * f + 1]}, outwardNorm(m, f, cc));
      const int bt = bcType ? bcType[f] : ORC_BC_NORMAL_GRADIENT;
      if (bt == ORC_BC_FIXED) {
        s[0] = flux * uFace[2 * f];
        s[1] = flux * uFace[2 * f + 1];
        addSrc(*e, m, 2, c, s);
      } else if (bt == ORC_BC_NORMAL_GRADIENT) {
        addS(*e, m, 2, c, c, flux);
        jdot(c, m.fC[f] - cc, flux, s);
        addSrc(*e, m, 2, c, s);
      } else if (bt == ORC_BC_SYMMETRY) {
      } else {
        setErr(err, errLen, "fv::div: unrecognized or unspecified boundary type.");
        delete e;
        return nullptr;
      }
    }
  }
  return e;
}

// VectorFiniteVolumeField::faceToCell (VectorFiniteVolumeField.cpp:5-26) and
// its weighted form (:28-51): per cell, sum over its faces of
// pointwise(v_f, |S_f|) [/ faceWeight_f], divided componentwise by sum |S_f|
// [times cellWeight]; null weights give the unweighted form
void orc_face_to_cell(const void *mesh, const double *faceVec, const double *cellWeight, const double *faceWeight,
                      double *out) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  for (int c = 0; c < m.nCells; ++c) {
    double sx = 0., sy = 0., tx = 0., ty = 0.;
    auto add = [&](int f) {
      const V2 n = m.fSf[f];
      const double ax = std::abs(n.x), ay = std::abs(n.y);
      double px = faceVec[2 * f] * ax, py = faceVec[2 * f + 1] * ay;
      if (faceWeight) {
        px /= faceWeight[f];
        py /= faceWeight[f];
      }
      tx += px;
      ty += py;
      sx += ax;
      sy += ay;
    };
    for (int k = m.iptr[c]; k < m.iptr[c + 1]; ++k) add(m.iface[k]);
    for (int k = m.bptr[c]; k < m.bptr[c + 1]; ++k) add(m.bface[k]);
    double vx = tx / sx, vy = ty / sy;
    if (cellWeight) {
      vx = cellWeight[c] * vx;
      vy = cellWeight[c] * vy;
    }
    out[2 * c] = vx;
    out[2 * c + 1] = vy;
  }
}

// ---------------------------------------------------------------- VOF (CICSAM)
namespace {
// std::min / std::max as the reference calls them -- min(a, b) = (b < a) ? b : a,
// max(a, b) = (a < b) ? b : a -- so a NaN first argument passes through
inline double smin(double a, double b) { return (b < a) ? b : a; }
inline double smax(double a, double b) { return (a < b) ? b : a; }
// clamp(v, lo, hi) = max(min(v, hi), lo), Math/Algorithm.h:11-13
inline double sclamp(double v, double lo, double hi) { return smax(smin(v, hi), lo); }
// cicsam::hc, Cicsam.cpp:5-8
inline double cicsamHc(double g, double co) { return g >= 0 && g <= 1 ? smin(1., g / co) : g; }
// cicsam::uq, Cicsam.cpp:10-18
inline double cicsamUq(double g, double co) {
  return g >= 0 && g <= 1 ? smin((8. * co * g + (1. - co) * (6. * g + 3.)) / 8., cicsamHc(g, co)) : g;
}
}  // namespace

double orc_cicsam_hc(double g, double co) { return cicsamHc(g, co); }
double orc_cicsam_uq(double g, double co) { return cicsamUq(g, co); }

// cicsam::faceInterpolationWeights, Cicsam.cpp:20-66 (k = 1; boundary faces 0).
// cos(2 theta_F) with theta_F = acos(c) is evaluated as 2 c^2 - 1, which is
// the same quantity without the two transcendental roundings; c > 1 from
// rounding (where acos is NaN) stays NaN, so the weight falls to 0 as there.
void orc_cicsam_weights(const void *mesh, const double *uFace, const double *gammaCell, const double *gradCell,
                        double timeStep, double *beta) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  for (int f = 0; f < m.nFaces; ++f) beta[f] = 0.;
  for (int f = 0; f < m.nFaces; ++f) {
    const int l = m.fL[f], r = m.fR[f];
    if (r < 0) continue;
    const V2 uf{uFace[2 * f], uFace[2 * f + 1]};
    const double flux = dot(uf, outwardNorm(m, f, m.cC[l]));
    const int d = flux > 0. ? l : r, a = flux <= 0. ? l : r;
    const V2 rc = m.cC[a] - m.cC[d];
    const V2 gd{gradCell[2 * d], gradCell[2 * d + 1]};
    const double gD = sclamp(gammaCell[d], 0., 1.), gA = sclamp(gammaCell[a], 0., 1.);
    const double gU = sclamp(gA - 2. * dot(rc, gd), 0., 1.);
    double gDT = (gD - gU) / (gA - gU);
    if (!std::isfinite(gDT)) gDT = 0.;
    // donor Courant number over all of the donor's faces
    double coD = 0.;
    for (int k = m.iptr[d]; k < m.iptr[d + 1]; ++k) {
      const int g = m.iface[k];
      coD += smax(dot(V2{uFace[2 * g], uFace[2 * g + 1]}, outwardNorm(m, g, m.cC[d])) / m.vol[d] * timeStep, 0.);
    }
    for (int k = m.bptr[d]; k < m.bptr[d + 1]; ++k) {
      const int g = m.bface[k];
      coD += smax(dot(V2{uFace[2 * g], uFace[2 * g + 1]}, outwardNorm(m, g, m.cC[d])) / m.vol[d] * timeStep, 0.);
    }
    const double mg = std::sqrt(gd.x * gd.x + gd.y * gd.y), mr = std::sqrt(rc.x * rc.x + rc.y * rc.y);
    const double c = std::abs(dot(V2{gd.x / mg, gd.y / mg}, V2{rc.x / mr, rc.y / mr}));
    const double cos2 = c > 1. ? NAN : 2. * c * c - 1.;
    const double psiF = smin(1. * (cos2 + 1.) / 2., 1.);
    const double gFT = psiF * cicsamHc(gDT, coD) + (1. - psiF) * cicsamUq(gDT, coD);
    const double b = (gFT - gDT) / (1. - gDT);
    beta[f] = std::isfinite(b) ? smax(smin(1., b), 0.) : 0.;
  }
}

// cicsam::div(u, gamma, beta, theta), Cicsam.cpp:87-138: per link the donor and
// acceptor coefficients theta (1 - b) F and theta b F in that order, the
// explicit part (1 - theta) F gamma0_f as a source; boundary links by type.
void *orc_cicsam_div(const void *mesh, const double *uFace, const double *beta, double theta, const int *bcType,
                     const double *gammaFace, const double *gamma0Cell, const double *gamma0Face, char *err,
                     int errLen) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  Eqn *e = new Eqn(m.nCells, 5);
  for (int c = 0; c < m.nCells; ++c) {
    const V2 cc = m.cC[c];
    for (int k = m.iptr[c]; k < m.iptr[c + 1]; ++k) {
      const int f = m.iface[k], nb = m.inb[k];
      const double flux = dot(V2{uFace[2 * f], uFace[2 * f + 1]}, outwardNorm(m, f, cc));
      const int d = flux > 0. ? c : nb, a = flux <= 0. ? c : nb;
      const double b = beta[f];
      e->addCoeff(c, d, theta * (1. - b) * flux);
      e->addCoeff(c, a, theta * b * flux);
      const double gF = (1. - b) * gamma0Cell[d] + b * gamma0Cell[a];
      e->rhs[c] += (1. - theta) * flux * gF;
    }
    for (int k = m.bptr[c]; k < m.bptr[c + 1]; ++k) {
      const int f = m.bface[k];
      const double flux = dot(V2{uFace[2 * f], uFace[2 * f + 1]}, outwardNorm(m, f, cc));
      const int bt = bcType ? bcType[f] : ORC_BC_NORMAL_GRADIENT;
      if (bt == ORC_BC_FIXED) {
        e->rhs[c] += theta * flux * gammaFace[f];
        e->rhs[c] += (1. - theta) * flux * gamma0Face[f];
      } else if (bt == ORC_BC_NORMAL_GRADIENT) {
        e->addCoeff(c, c, theta * flux);
        e->rhs[c] += (1. - theta) * flux * gamma0Cell[c];
      } else if (bt == ORC_BC_SYMMETRY) {
      } else {
        setErr(err, errLen, "cicsam::div: unrecognized or unspecified boundary type.");
        delete e;
        return nullptr;
      }
    }
  }
  return e;
}

// cicsam::computeMomentumFlux, Cicsam.cpp:68-85: rho(g) u_f with the same face
// value g the gamma equation used (interior) or the face value (boundary)
void orc_cicsam_momentum_flux(const void *mesh, double rho1, double rho2, const double *uFace,
                              const double *gammaCell, const double *gammaFace, const double *beta, double *rhoU) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  for (int f = 0; f < m.nFaces; ++f) {
    const int l = m.fL[f], r = m.fR[f];
    double g;
    if (r >= 0) {
      const double flux = dot(V2{uFace[2 * f], uFace[2 * f + 1]}, outwardNorm(m, f, m.cC[l]));
      const int d = flux > 0. ? l : r, a = flux <= 0. ? l : r;
      const double b = beta[f];
      g = (1. - b) * gammaCell[d] + b * gammaCell[a];
    } else {
      g = gammaFace[f];
    }
    const double rho = rho1 + sclamp(g, 0., 1.) * (rho2 - rho1);
    rhoU[2 * f] = rho * uFace[2 * f];
    rhoU[2 * f + 1] = rho * uFace[2 * f + 1];
  }
}

// FractionalStepMultiphase::updateProperties, FractionalStepMultiphase.cpp:175-214:
// density and (from the kinematic viscosities) dynamic viscosity on cells and faces
void orc_vof_properties(const void *mesh, double rho1, double rho2, double mu1, double mu2,
                        const double *gammaCell, const double *gammaFace, double *rhoCell, double *rhoFace,
                        double *muCell, double *muFace) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  for (int c = 0; c < m.nCells; ++c) {
    rhoCell[c] = rho1 + sclamp(gammaCell[c], 0., 1.) * (rho2 - rho1);
    muCell[c] = rhoCell[c] / (rho1 / mu1 + sclamp(gammaCell[c], 0., 1.) * (rho2 / mu2 - rho1 / mu1));
  }
  for (int f = 0; f < m.nFaces; ++f) {
    rhoFace[f] = rho1 + sclamp(gammaFace[f], 0., 1.) * (rho2 - rho1);
    muFace[f] = rhoFace[f] / (rho1 / mu1 + sclamp(gammaFace[f], 0., 1.) * (rho2 / mu2 - rho1 / mu1));
  }
}

// src::div(u), Source.cpp:5-21
void orc_src_div(const void *mesh, const double *uFace, double *out) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  for (int c = 0; c < m.nCells; ++c) {
    double d = 0.;
    for (int k = m.iptr[c]; k < m.iptr[c + 1]; ++k) {
      int f = m.iface[k];
      d += dot(V2{uFace[2 * f], uFace[2 * f + 1]}, outwardNorm(m, f, m.cC[c]));
    }
    for (int k = m.bptr[c]; k < m.bptr[c + 1]; ++k) {
      int f = m.bface[k];
      d += dot(V2{uFace[2 * f], uFace[2 * f + 1]}, outwardNorm(m, f, m.cC[c]));
    }
    out[c] = d;
  }
}

// src::src(field), Source.cpp:77-95 (block layout for vectors)
void orc_src_src(const void *mesh, int ncomp, const double *phiCell, double *out) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  for (int c = 0; c < m.nCells; ++c)
    for (int q = 0; q < ncomp; ++q) out[rowOf(m, c, q)] = phiCell[ncomp * c + q] * m.vol[c];
}

// ScalarGradient::computeFaces + compute(FACE_TO_CELL), ScalarGradient.cpp:34-74
void orc_grad(const void *mesh, const double *phiCell, const double *phiFace, double *gradFace,
              double *gradCell) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  for (int f = 0; f < m.nFaces; ++f) {
    int l = m.fL[f], r = m.fR[f];
    V2 d;
    double dphi;
    if (r >= 0) { d = m.cC[r] - m.cC[l]; dphi = phiCell[r] - phiCell[l]; }
    else { d = m.fC[f] - m.cC[l]; dphi = phiFace[f] - phiCell[l]; }
    double ms = magSqr(d);
    gradFace[2 * f] = d.x * dphi / ms;
    gradFace[2 * f + 1] = d.y * dphi / ms;
  }
  for (int c = 0; c < m.nCells; ++c) {
    double sx = 0., sy = 0., tx = 0., ty = 0.;
    auto acc = [&](int f) {
      V2 n = outwardNorm(m, f, m.cC[c]);
      double ax = std::fabs(n.x), ay = std::fabs(n.y);
      tx += gradFace[2 * f] * ax;
      ty += gradFace[2 * f + 1] * ay;
      sx += ax;
      sy += ay;
    };
    for (int k = m.iptr[c]; k < m.iptr[c + 1]; ++k) acc(m.iface[k]);
    for (int k = m.bptr[c]; k < m.bptr[c + 1]; ++k) acc(m.bface[k]);
    gradCell[2 * c] = tx / sx;
    gradCell[2 * c + 1] = ty / sy;
  }
}

// FiniteVolumeField<T>::interpolateFaces(DISTANCE) + setBoundaryFaces
// (FiniteVolumeField.tpp:129-182, VectorFiniteVolumeField.cpp:140-160).
// phiFace is in/out: FIXED faces keep their value.
void orc_interpolate_faces(const void *mesh, int ncomp, const double *phiCell, const int *bcType,
                           double *phiFace) {
  const Mesh &m = *static_cast<const Mesh *>(mesh);
  for (int f = 0; f < m.nFaces; ++f) {
    int l = m.fL[f], r = m.fR[f];
    if (r >= 0) {
      double l1 = mag(m.fC[f] - m.cC[r]);
      double l2 = mag(m.fC[f] - m.cC[l]);
      double g = l1 / (l1 + l2);  // Face::distanceWeight, Face.cpp:60-64
      for (int q = 0; q < ncomp; ++q)
        phiFace[ncomp * f + q] = g * phiCell[ncomp * l + q] + (1. - g) * phiCell[ncomp * r + q];
    } else {
      int bt = bcType ? bcType[f] : ORC_BC_NORMAL_GRADIENT;
      if (bt == ORC_BC_FIXED) continue;
      if (bt == ORC_BC_SYMMETRY && ncomp == 2) {
        V2 n = m.fSf[f];
        double px = phiCell[2 * l], py = phiCell[2 * l + 1];
        double d = px * n.x + py * n.y;
        double ms = magSqr(n);
        phiFace[2 * f] = px - n.x * d / ms;
        phiFace[2 * f + 1] = py - n.y * d / ms;
      } else {
        for (int q = 0; q < ncomp; ++q) phiFace[ncomp * f + q] = phiCell[ncomp * l + q];
      }
    }
  }
}

// ---------------------------------------------------------------- algebra
// y = A x over the non-empty slots, in storage order, fused multiply-add.
void orc_spmv(int n, const int *rowPtr, const int *colInd, const double *vals, const double *x,
              double *y) {
  for (int i = 0; i < n; ++i) {
    double acc = 0.;
    for (int j = rowPtr[i]; j < rowPtr[i + 1]; ++j)
      if (colInd[j] >= 0) acc = std::fma(vals[j], x[colInd[j]], acc);
    y[i] = acc;
  }
}

// ILU(0), IKJ variant (Saad, Iterative Methods for Sparse Linear Systems, 2nd
// ed., Alg. 10.4): in place on a compact CSR whose rows are sorted by column
// and contain the diagonal.  L is unit lower (strict part stored), U holds the
// diagonal.  This is the RILUK(level 0) factor Ifpack2 computes inside the
// reference's Schwarz preconditioner (TrilinosBelosSparseMatrixSolver.cpp:69-82).
// Returns 0, or -(row+1) for a missing or zero pivot.
int orc_ilu0(int n, const int *rowPtr, const int *colInd, double *vals) {
  std::vector<int> diag(n, -1), pos(n, -1);
  for (int i = 0; i < n; ++i)
    for (int j = rowPtr[i]; j < rowPtr[i + 1]; ++j)
      if (colInd[j] == i) diag[i] = j;
  for (int i = 0; i < n; ++i) {
    if (diag[i] < 0) return -(i + 1);
    for (int j = rowPtr[i]; j < rowPtr[i + 1]; ++j) pos[colInd[j]] = j;
    for (int jj = rowPtr[i]; jj < rowPtr[i + 1] && colInd[jj] < i; ++jj) {
      int k = colInd[jj];
      double lik = vals[jj] / vals[diag[k]];
      vals[jj] = lik;
      for (int kk = diag[k] + 1; kk < rowPtr[k + 1]; ++kk) {
        int p = pos[colInd[kk]];
        if (p >= 0) vals[p] = std::fma(-lik, vals[kk], vals[p]);
      }
    }
    if (vals[diag[i]] == 0.) return -(i + 1);
    for (int j = rowPtr[i]; j < rowPtr[i + 1]; ++j) pos[colInd[j]] = -1;
  }
  return 0;
}

}  // extern "C"

namespace {

struct Csr {
  int n = 0;
  std::vector<int> rp, ci;
  std::vector<double> v;
};

// Compact a padded reference CSR (drop -1 slots) and sort each row by column.
Csr compactSorted(int n, const int *rowPtr, const int *colInd, const double *vals) {
  Csr c;
  c.n = n;
  c.rp.assign(n + 1, 0);
  std::vector<std::pair<int, double>> row;
  for (int i = 0; i < n; ++i) {
    row.clear();
    for (int j = rowPtr[i]; j < rowPtr[i + 1]; ++j)
      if (colInd[j] >= 0) row.push_back({colInd[j], vals[j]});
    std::sort(row.begin(), row.end(),
              [](const std::pair<int, double> &a, const std::pair<int, double> &b) {
                return a.first < b.first;
              });
    for (auto &e : row) { c.ci.push_back(e.first); c.v.push_back(e.second); }
    c.rp[i + 1] = (int)c.ci.size();
  }
  return c;
}

// Forward/backward substitution with an LU factor stored in one sorted CSR.
void luSolve(const Csr &f, const std::vector<int> &diag, const double *b, double *x) {
  for (int i = 0; i < f.n; ++i) {
    double acc = b[i];
    for (int j = f.rp[i]; j < diag[i]; ++j) acc = std::fma(-f.v[j], x[f.ci[j]], acc);
    x[i] = acc;
  }
  for (int i = f.n - 1; i >= 0; --i) {
    double acc = x[i];
    for (int j = diag[i] + 1; j < f.rp[i + 1]; ++j) acc = std::fma(-f.v[j], x[f.ci[j]], acc);
    x[i] = acc / f.v[diag[i]];
  }
}

double dotv(const std::vector<double> &a, const std::vector<double> &b) {
  double s = 0.;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

// Right-preconditioned BiCGStab (van der Vorst 1992), the method Belos runs
// for "BICGSTAB" with Ifpack2 set as right preconditioner
// (TrilinosBelosSparseMatrixSolver.cpp:18-42).  Stops on the FP64 true
// residual ||b - A x||_2 / ||b||_2 <= tol (recomputed whenever the recursive
// residual says converged).
template <class Prec>
int bicgstab(const Csr &A, const double *b, double *x, double tol, int maxIters, Prec prec,
             int *itersOut, double *relres) {
  int n = A.n;
  std::vector<double> r(n), r0(n), p(n, 0.), v(n, 0.), s(n), t(n), ph(n), sh(n);
  auto matvec = [&](const double *in, double *out) {
    for (int i = 0; i < n; ++i) {
      double acc = 0.;
      for (int j = A.rp[i]; j < A.rp[i + 1]; ++j) acc = std::fma(A.v[j], in[A.ci[j]], acc);
      out[i] = acc;
    }
  };
  std::vector<double> bv(b, b + n);
  double bnorm = std::sqrt(dotv(bv, bv));
  if (bnorm == 0.) {
    std::fill(x, x + n, 0.);
    *itersOut = 0;
    *relres = 0.;
    return 0;
  }
  auto trueRes = [&]() {
    matvec(x, t.data());
    for (int i = 0; i < n; ++i) r[i] = b[i] - t[i];
    return std::sqrt(dotv(r, r)) / bnorm;
  };
  double res = trueRes();
  int it = 0;
  while (res > tol && it < maxIters) {
    r0 = r;
    double rho = 1., alpha = 1., omega = 1.;
    std::fill(p.begin(), p.end(), 0.);
    std::fill(v.begin(), v.end(), 0.);
    bool recConv = false;
    while (it < maxIters) {
      ++it;
      double rhoNew = dotv(r0, r);
      if (rhoNew == 0.) break;
      double beta = (rhoNew / rho) * (alpha / omega);
      for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
      prec(p.data(), ph.data());
      matvec(ph.data(), v.data());
      alpha = rhoNew / dotv(r0, v);
      for (int i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
      prec(s.data(), sh.data());
      matvec(sh.data(), t.data());
      double tt = dotv(t, t);
      omega = tt > 0. ? dotv(t, s) / tt : 0.;
      for (int i = 0; i < n; ++i) x[i] += alpha * ph[i] + omega * sh[i];
      for (int i = 0; i < n; ++i) r[i] = s[i] - omega * t[i];
      rho = rhoNew;
      if (std::sqrt(dotv(r, r)) / bnorm <= tol) { recConv = true; break; }
      if (omega == 0.) break;
    }
    res = trueRes();
    if (!recConv && res > tol && it >= maxIters) break;
  }
  *itersOut = it;
  *relres = res;
  return res <= tol ? 0 : 1;
}

}  // namespace

extern "C" {

// The same BiCGStab + ILU(0) with the SpMV, dot products and vector updates
// shared over OpenMP threads (the triangular ILU(0) sweeps stay sequential):
// the all-cores CPU timing baseline.  Reductions are split over threads, so
// results differ from the sequential solver in the last bits; the sequential
// one is the checker.
int orc_bicgstab_ilu0_omp(int n, const int *rowPtr, const int *colInd, const double *vals, const double *b,
                          double *x, double tol, int maxIters, int nThreads, int *itersOut, double *relres) {
  Csr A = compactSorted(n, rowPtr, colInd, vals);
  Csr F = A;
  if (orc_ilu0(n, F.rp.data(), F.ci.data(), F.v.data()) != 0) return -1;
  std::vector<int> diag(n);
  for (int i = 0; i < n; ++i)
    for (int j = F.rp[i]; j < F.rp[i + 1]; ++j)
      if (F.ci[j] == i) diag[i] = j;
  const int nt = nThreads > 0 ? nThreads : 1;
  std::vector<double> r(n), r0(n), p(n, 0.), v(n, 0.), s(n), t(n), ph(n), sh(n);
  auto matvec = [&](const double *in, double *out) {
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int i = 0; i < n; ++i) {
      double acc = 0.;
      for (int j = A.rp[i]; j < A.rp[i + 1]; ++j) acc = std::fma(A.v[j], in[A.ci[j]], acc);
      out[i] = acc;
    }
  };
  auto dot = [&](const double *a, const double *c) {
    double sum = 0.;
#pragma omp parallel for num_threads(nt) schedule(static) reduction(+ : sum)
    for (int i = 0; i < n; ++i) sum += a[i] * c[i];
    return sum;
  };
  double bnorm = std::sqrt(dot(b, b));
  if (bnorm == 0.) {
    std::fill(x, x + n, 0.);
    *itersOut = 0;
    *relres = 0.;
    return 0;
  }
  auto trueRes = [&]() {
    matvec(x, t.data());
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int i = 0; i < n; ++i) r[i] = b[i] - t[i];
    return std::sqrt(dot(r.data(), r.data())) / bnorm;
  };
  double res = trueRes();
  int it = 0;
  while (res > tol && it < maxIters) {
    r0 = r;
    double rho = 1., alpha = 1., omega = 1.;
    std::fill(p.begin(), p.end(), 0.);
    std::fill(v.begin(), v.end(), 0.);
    bool recConv = false;
    while (it < maxIters) {
      ++it;
      const double rhoNew = dot(r0.data(), r.data());
      if (rhoNew == 0.) break;
      const double beta = (rhoNew / rho) * (alpha / omega);
#pragma omp parallel for num_threads(nt) schedule(static)
      for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
      luSolve(F, diag, p.data(), ph.data());
      matvec(ph.data(), v.data());
      alpha = rhoNew / dot(r0.data(), v.data());
#pragma omp parallel for num_threads(nt) schedule(static)
      for (int i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
      luSolve(F, diag, s.data(), sh.data());
      matvec(sh.data(), t.data());
      const double tt = dot(t.data(), t.data());
      omega = tt > 0. ? dot(t.data(), s.data()) / tt : 0.;
#pragma omp parallel for num_threads(nt) schedule(static)
      for (int i = 0; i < n; ++i) {
        x[i] += alpha * ph[i] + omega * sh[i];
        r[i] = s[i] - omega * t[i];
      }
      rho = rhoNew;
      if (std::sqrt(dot(r.data(), r.data())) / bnorm <= tol) { recConv = true; break; }
      if (omega == 0.) break;
    }
    res = trueRes();
    if (!recConv && res > tol && it >= maxIters) break;
  }
  *itersOut = it;
  *relres = res;
  return res <= tol ? 0 : 1;
}

int orc_bicgstab_ilu0(int n, const int *rowPtr, const int *colInd, const double *vals,
                      const double *b, double *x, double tol, int maxIters, int *iters,
                      double *relres) {
  Csr A = compactSorted(n, rowPtr, colInd, vals);
  Csr F = A;
  int rc = orc_ilu0(n, F.rp.data(), F.ci.data(), F.v.data());
  if (rc != 0) return -1;
  std::vector<int> diag(n);
  for (int i = 0; i < n; ++i)
    for (int j = F.rp[i]; j < F.rp[i + 1]; ++j)
      if (F.ci[j] == i) diag[i] = j;
  auto prec = [&](const double *in, double *out) { luSolve(F, diag, in, out); };
  return bicgstab(A, b, x, tol, maxIters, prec, iters, relres);
}

// ILUT(tau, p) (Saad, Alg. 10.6) with the parameters of Eigen's IncompleteLUT
// (drop tolerance relative to the row 2-norm; at most p = fill entries kept in
// each of the L and U parts of a row, p = fillFactor * nnz(A) / n + 1), in the
// natural order.  Eigen additionally applies an AMD pre-ordering, which this
// restatement omits.
int orc_ilut(int n, const int *rowPtr, const int *colInd, const double *vals, double dropTol,
             int fillFactor, int **outRowPtr, int **outColInd, double **outVals) {
  Csr A = compactSorted(n, rowPtr, colInd, vals);
  long nnzA = A.rp[n];
  int p = (int)((nnzA * (long)fillFactor) / n) + 1;
  std::vector<int> frp(n + 1, 0), fci, fdiag(n);
  std::vector<double> fv;
  std::vector<double> w(n, 0.);
  std::vector<int> inw(n, 0), lidx, uidx;
  for (int i = 0; i < n; ++i) {
    double rn = 0.;
    lidx.clear();
    uidx.clear();
    bool hasDiag = false;
    for (int j = A.rp[i]; j < A.rp[i + 1]; ++j) {
      int c = A.ci[j];
      w[c] = A.v[j];
      inw[c] = 1;
      rn += A.v[j] * A.v[j];
      if (c < i) lidx.push_back(c); else if (c > i) uidx.push_back(c); else hasDiag = true;
    }
    if (!hasDiag) { w[i] = 0.; inw[i] = 1; }
    double tau = dropTol * std::sqrt(rn);
    // eliminate in increasing column order; fill may add new L columns
    std::sort(lidx.begin(), lidx.end());
    std::vector<int> keptL;
    size_t q = 0;
    while (q < lidx.size()) {
      int k = lidx[q++];
      double wk = w[k] / fv[fdiag[k]];
      if (std::fabs(wk) < tau) { w[k] = 0.; inw[k] = 0; continue; }
      w[k] = wk;
      keptL.push_back(k);
      for (int kk = fdiag[k] + 1; kk < frp[k + 1]; ++kk) {
        int c = fci[kk];
        if (!inw[c]) {
          inw[c] = 1;
          w[c] = 0.;
          if (c < i) {
            lidx.insert(std::upper_bound(lidx.begin() + q, lidx.end(), c), c);
          } else if (c > i) {
            uidx.push_back(c);
          }
        }
        w[c] = std::fma(-wk, fv[kk], w[c]);
      }
    }
    auto keepLargest = [&](std::vector<int> &idx) {
      std::vector<int> out;
      for (int c : idx)
        if (std::fabs(w[c]) >= tau) out.push_back(c);
      if ((int)out.size() > p) {
        std::nth_element(out.begin(), out.begin() + p, out.end(),
                         [&](int a, int b) { return std::fabs(w[a]) > std::fabs(w[b]); });
        out.resize(p);
      }
      std::sort(out.begin(), out.end());
      return out;
    };
    std::vector<int> L = keepLargest(keptL), U = keepLargest(uidx);
    for (int c : L) { fci.push_back(c); fv.push_back(w[c]); }
    fdiag[i] = (int)fci.size();
    double d = w[i];
    if (d == 0.) d = tau > 0. ? tau : 1.;  // replace a zero pivot
    fci.push_back(i);
    fv.push_back(d);
    for (int c : U) { fci.push_back(c); fv.push_back(w[c]); }
    frp[i + 1] = (int)fci.size();
    for (int c : lidx) { w[c] = 0.; inw[c] = 0; }
    for (int c : uidx) { w[c] = 0.; inw[c] = 0; }
    w[i] = 0.;
    inw[i] = 0;
  }
  *outRowPtr = (int *)std::malloc(sizeof(int) * (n + 1));
  *outColInd = (int *)std::malloc(sizeof(int) * fci.size());
  *outVals = (double *)std::malloc(sizeof(double) * fv.size());
  std::copy(frp.begin(), frp.end(), *outRowPtr);
  std::copy(fci.begin(), fci.end(), *outColInd);
  std::copy(fv.begin(), fv.end(), *outVals);
  return 0;
}

int orc_bicgstab_ilut(int n, const int *rowPtr, const int *colInd, const double *vals,
                      const double *b, double *x, double tol, int maxIters, double dropTol,
                      int fillFactor, int *iters, double *relres) {
  Csr A = compactSorted(n, rowPtr, colInd, vals);
  int *frp, *fci;
  double *fv;
  orc_ilut(n, rowPtr, colInd, vals, dropTol, fillFactor, &frp, &fci, &fv);
  Csr F;
  F.n = n;
  F.rp.assign(frp, frp + n + 1);
  F.ci.assign(fci, fci + frp[n]);
  F.v.assign(fv, fv + frp[n]);
  std::free(frp); std::free(fci); std::free(fv);
  std::vector<int> diag(n);
  for (int i = 0; i < n; ++i)
    for (int j = F.rp[i]; j < F.rp[i + 1]; ++j)
      if (F.ci[j] == i) diag[i] = j;
  auto prec = [&](const double *in, double *out) { luSolve(F, diag, in, out); };
  return bicgstab(A, b, x, tol, maxIters, prec, iters, relres);
}

// Forward (unit L) then backward (U) substitution with an ILU factor held in
// one column-sorted CSR (L strictly below, U on and above the diagonal), in
// ascending column order per row: y = L^{-1} b, x = U^{-1} y.
void orc_lu_solve(int n, const int *rowPtr, const int *colInd, const double *vals, const double *b,
                  double *y, double *x) {
  for (int i = 0; i < n; ++i) {
    double acc = b[i];
    for (int j = rowPtr[i]; j < rowPtr[i + 1] && colInd[j] < i; ++j) acc = std::fma(-vals[j], y[colInd[j]], acc);
    y[i] = acc;
  }
  for (int i = n - 1; i >= 0; --i) {
    double acc = y[i], d = 0.;
    for (int j = rowPtr[i]; j < rowPtr[i + 1]; ++j) {
      if (colInd[j] == i) d = vals[j];
      else if (colInd[j] > i) acc = std::fma(-vals[j], x[colInd[j]], acc);
    }
    x[i] = acc / d;
  }
}

void orc_free(void *p) { std::free(p); }

}  // extern "C"
