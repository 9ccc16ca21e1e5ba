"""TEST INFRASTRUCTURE ONLY.  The ghost-cell immersed-boundary rows
(src/2D/Unstructured/FiniteVolume/ImmersedBoundary/GhostCellImmersedBoundary.cpp,
GhostCellImmersedBoundaryBcStencil.cpp, ...FixedBcStencil.cpp,
...NormalGradientBcStencil.cpp) for circular objects, restated on the CPU
with numpy and the oracle's CrsEquation; the checker for the device rows.
Only tests/ may use it.

Cells (GhostCellImmersedBoundary::updateCells, :12-31): per object, in
object order, the domain cells whose centroid lies strictly inside the circle
are solid; those with a face or diagonal (node) neighbour outside the circle
(ImmersedBoundaryObject::internalPerimeterCells with diagonals,
ImmersedBoundaryObject.tpp:5-35) are the ghost (ib) cells; both leave the
fluid cells.

Stencil of a ghost cell c (BcStencil, GhostCellImmersedBoundaryBcStencil.cpp:3-70):
boundary point bp = centre + R (x_c - centre) / |x_c - centre|
(Circle::nearestIntersect), image point ip = 2 bp - x_c, wall normal
nw = (centre - bp) / |centre - bp|; the four cells of the mesh node nearest
to ip (FiniteVolumeGrid2D::findNearestNode, lowest node id on a tie; the
reference takes its R-tree's answer) with the bilinear interpolation
Ainv = inverse of the rows (x y, x, y, 1) of their centroids.
  FIXED (FixedBcStencil.cpp:17-22): coefficients (ip.x ip.y, ip.x, ip.y, 1) Ainv / 2
      on the four cells and 1/2 on c, source -refValue
  NORMAL_GRADIENT (NormalGradientBcStencil.cpp:22-42): (bp.y nw.x + bp.x nw.y,
      nw.x, nw.y, 0) Ainv when c is one of the four cells, else
      -(ip.x ip.y, ip.x, ip.y, 1) Ainv / l on the four and 1 / l on c,
      l = |ip - x_c|; source -refValue
Solid cells: a unit diagonal with source -refValue (FIXED) or 0 (NORMAL_GRADIENT).
Velocity rows (velocityBcs, :73-90): FIXED stencils on both components with
source -(object velocity), every object here translating (no rotation).

Ainv is numpy's inverse where the reference calls LAPACK getrf/getri: the
device rows agree to rounding.
"""
import numpy as np

from oracle import oracle as O

FLUID, GHOST, SOLID = 0, 1, 2


class GhostCellOracle:
    def __init__(self, mesh, xy, cptr, cind, circles):
        self.m = mesh
        N = mesh.n_cells
        xy = np.asarray(xy, float).reshape(-1, 2)
        cptr, cind = np.asarray(cptr), np.asarray(cind)
        self.xy = xy
        _, self.cc = mesh.cells()
        L = mesh.links()
        iptr, inb = L["iptr"], L["inb"]
        self.node_cells = [[] for _ in range(len(xy))]
        for c in range(N):
            for k in range(cptr[c], cptr[c + 1]):
                self.node_cells[cind[k]].append(c)
        # cell links: face neighbours, then diagonal (node) neighbours
        links = []
        for c in range(N):
            nbs = list(inb[iptr[c]:iptr[c + 1]])
            seen = set(nbs)
            for k in range(cptr[c], cptr[c + 1]):
                for kc in self.node_cells[cind[k]]:
                    if kc != c and kc not in seen:
                        nbs.append(kc)
                        seen.add(kc)
            links.append(nbs)
        self.types = np.zeros(N, np.int32)
        self.owner = np.full(N, -1, np.int32)
        self.circles = [tuple(map(float, c)) for c in circles]
        for o, (cx, cy, r) in enumerate(self.circles):
            d = self.cc - np.array([cx, cy])
            inside = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] < r * r  # Circle::isInside
            for c in np.where(inside & (self.types == FLUID))[0]:
                ghost = any(not inside[k] for k in links[c])
                self.types[c] = GHOST if ghost else SOLID
                self.owner[c] = o
        self.stencils = {}
        for c in np.where(self.types == GHOST)[0]:
            self.stencils[int(c)] = self._stencil(int(c), self.circles[self.owner[c]])

    def _stencil(self, c, circle):
        cx, cy, r = circle
        xc = self.cc[c]
        centre = np.array([cx, cy])
        d = xc - centre
        bp = d / np.sqrt(d[0] * d[0] + d[1] * d[1]) * r + centre
        ip = 2. * bp - xc
        w = centre - bp
        nw = w / np.sqrt(w[0] * w[0] + w[1] * w[1])
        dn = self.xy - ip
        node = int(np.argmin(dn[:, 0] * dn[:, 0] + dn[:, 1] * dn[:, 1]))
        cells = list(self.node_cells[node])
        if len(cells) != 4:
            raise ValueError("ghost-cell stencil of cell %d: nearest node has %d cells" % (c, len(cells)))
        x = self.cc[cells]
        A = np.c_[x[:, 0] * x[:, 1], x[:, 0], x[:, 1], np.ones(4)]
        Ainv = np.linalg.inv(A)
        v = np.array([ip[0] * ip[1], ip[0], ip[1], 1.])
        fixed = (list(cells) + [c], list(v @ Ainv / 2.) + [0.5])
        if c in cells:
            g = np.array([bp[1] * nw[0] + bp[0] * nw[1], nw[0], nw[1], 0.])
            ng = (list(cells), list(g @ Ainv))
        else:
            e = ip - xc
            l = np.sqrt(e[0] * e[0] + e[1] * e[1])
            ng = (list(cells) + [c], list(-(v @ Ainv) / l) + [1. / l])
        return dict(bp=bp, ip=ip, nw=nw, cells=cells, fixed=fixed, ng=ng)

    def bcs(self, bc_type, ref_value):
        """GhostCellImmersedBoundary::bcs: a fresh scalar equation holding the
        ghost and solid rows."""
        N = self.m.n_cells
        e = O.Eqn.create(N, 5)
        for c in range(N):
            if self.types[c] == GHOST:
                cols, vals = self.stencils[c]["fixed" if bc_type == O.BC_FIXED else "ng"]
                for k, v in zip(cols, vals):
                    e.add_coeff(c, k, v)
                e.add_rhs(c, -ref_value)
            elif self.types[c] == SOLID:
                e.add_coeff(c, c, 1.)
                if bc_type == O.BC_FIXED:
                    e.add_rhs(c, -ref_value)
        return e

    def velocity_bcs(self, velocities):
        """GhostCellImmersedBoundary::velocityBcs for translating objects
        (velocity per object): Vector2D rows in the block layout."""
        N = self.m.n_cells
        e = O.Eqn.create(2 * N, 5)
        for c in range(N):
            if self.types[c] == FLUID:
                continue
            u = velocities[self.owner[c]]
            if self.types[c] == GHOST:
                cols, vals = self.stencils[c]["fixed"]
                for q in range(2):
                    for k, v in zip(cols, vals):
                        e.add_coeff(q * N + c, q * N + k, v)
            else:
                for q in range(2):
                    e.add_coeff(q * N + c, q * N + c, 1.)
            for q in range(2):
                e.add_rhs(q * N + c, -u[q])
        return e
