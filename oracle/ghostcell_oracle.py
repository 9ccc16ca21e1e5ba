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
device rows agree to rounding.  Bit-level parity with the reference itself:
parity unpinned (LAPACK and the Boost R-tree of the nearest-node query are
absent); tests/test_oracle_ghostcell.py pins the rows by exactness on linear
fields, the convergence of a Poisson problem around a hole and the closed
forms of the body force.

Force on each body (applyHydrodynamicForce, GhostCellImmersedBoundary.cpp:94-146):
per ghost cell, the boundary pressure bpValue(p) = (p_c + p(ip)) / 2 plus
rho dot(bp, g), and the wall shear mu dot(dot(grad u(bp), nw), nw.tangentVec())
with grad u from the bilinear fit over the four cells of the node nearest to
bp (BcStencil::bpGrad); the points sorted by their angle about the centre;
force = sum over consecutive pairs of -(p_a + p_b)/2 (x_b - x_a).normalVec()
+ (s_a + s_b)/2 (x_b - x_a).
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

    def _node_cells(self, pt):
        dn = self.xy - pt
        node = int(np.argmin(dn[:, 0] * dn[:, 0] + dn[:, 1] * dn[:, 1]))
        return list(self.node_cells[node])

    def hydrodynamic_force(self, rho, mu, u_cell, p_cell, g=(0.0, 0.0)):
        """Force (fx, fy) per body; also returns the per-point (bp, pressure, shear)."""
        u = np.asarray(u_cell, float).reshape(-1, 2)
        p = np.asarray(p_cell, float)
        out, points = [], []
        for o, (cx, cy, r) in enumerate(self.circles):
            pts = []
            for c in np.where((self.types == GHOST) & (self.owner == o))[0]:
                st = self.stencils[int(c)]
                bp, ip, nw = st["bp"], st["ip"], st["nw"]
                cells = st["cells"]
                x = self.cc[cells]
                Ainv = np.linalg.inv(np.c_[x[:, 0] * x[:, 1], x[:, 0], x[:, 1], np.ones(4)])
                v = np.array([ip[0] * ip[1], ip[0], ip[1], 1.])
                pb = (p[c] + (v @ Ainv) @ p[cells]) / 2.
                pr = pb + rho * (bp[0] * g[0] + bp[1] * g[1])
                bc = self._node_cells(bp)
                y = self.cc[bc]
                X = np.linalg.solve(np.c_[y[:, 0] * y[:, 1], y[:, 0], y[:, 1], np.ones(4)], u[bc])
                G = np.array([[bp[1], 1., 0., 0.], [bp[0], 0., 1., 0.]]) @ X  # rows d/dx, d/dy
                T = np.array([[G[0, 0], G[0, 1]], [G[1, 0], G[1, 1]]])  # xx, xy / yx, yy
                tn = T @ nw
                sh = mu * (tn[0] * -nw[1] + tn[1] * nw[0])
                pts.append((np.arctan2(bp[1] - cy, bp[0] - cx), bp, pr, sh))
            pts.sort(key=lambda t: t[0])
            f = np.zeros(2)
            for i in range(len(pts)):
                _, pa, pra, sha = pts[i]
                _, pb2, prb, shb = pts[(i + 1) % len(pts)]
                d = pb2 - pa
                f += -(pra + prb) / 2. * np.array([d[1], -d[0]]) + (sha + shb) / 2. * d
            out.append(f)
            points.append(pts)
        return np.array(out), points

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
