"""TEST INFRASTRUCTURE ONLY.  The ghost-cell immersed-boundary rows
(src/2D/Unstructured/FiniteVolume/ImmersedBoundary/GhostCellImmersedBoundary.cpp,
GhostCellImmersedBoundaryBcStencil.cpp, ...FixedBcStencil.cpp,
...NormalGradientBcStencil.cpp) for circular and box objects, restated on the CPU
with numpy and the oracle's CrsEquation; the checker for the device rows.
Only tests/ may use it.

Cells (GhostCellImmersedBoundary::updateCells, :12-31): per object, in
object order, the domain cells whose centroid lies strictly inside the shape
are solid; those with a face or diagonal (node) neighbour outside the shape
(ImmersedBoundaryObject::internalPerimeterCells with diagonals,
ImmersedBoundaryObject.tpp:5-35) are the ghost (ib) cells; both leave the
fluid cells.

Stencil of a ghost cell c (BcStencil, GhostCellImmersedBoundaryBcStencil.cpp:3-70):
boundary point bp = nearestIntersect(x_c) (circle: centre + R (x_c - centre) /
|x_c - centre|; box: the nearest point of its edges, Box.cpp), image point
ip = 2 bp - x_c, wall normal nw = nearestEdgeUnitNormal(bp) (circle:
(centre - bp) / |centre - bp|; box: the unit normal of the nearest edge, turned
towards the centroid); the four cells of the mesh node nearest
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
+ (s_a + s_b)/2 (x_b - x_a).  Each body integrates its own points only: the
reference keeps one stress list across the body loop (declared before it,
never cleared, :100-108), so with two or more bodies its force on body k also
integrates the points of bodies 0..k-1.  That is a reference bug, fixed here
and on the device on purpose; the two-body closed form in
tests/test_oracle_ghostcell.py pins the per-body force.
"""
import numpy as np

from oracle import oracle as O

FLUID, GHOST, SOLID = 0, 1, 2


class GhostCellOracle:
    def __init__(self, mesh, xy, cptr, cind, circles, boxes=()):
        """circles: (cx, cy, r) each; boxes: (lower x, lower y, upper x, upper
        y) each (Box.cpp); bodies are the circles, then the boxes."""
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
        self.bodies = [("circle",) + c for c in self.circles] + [("box",) + tuple(map(float, b)) for b in boxes]
        for o, body in enumerate(self.bodies):
            inside = self._inside(body, self.cc)
            for c in np.where(inside & (self.types == FLUID))[0]:
                ghost = any(not inside[k] for k in links[c])
                self.types[c] = GHOST if ghost else SOLID
                self.owner[c] = o
        self.stencils = {}
        for c in np.where(self.types == GHOST)[0]:
            self.stencils[int(c)] = self._stencil(int(c), self.bodies[self.owner[c]])

    # shape queries: Circle.cpp (isInside, nearestIntersect), Box.cpp
    # (isInside, nearestIntersect, nearestEdge, edges), and
    # ImmersedBoundaryObject::nearestEdgeUnitNormal (ImmersedBoundaryObject.cpp:76-90)
    @staticmethod
    def _inside(body, pts):
        if body[0] == "circle":
            d = pts - np.array(body[1:3])
            return d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] < body[3] * body[3]
        lx, ly, ux, uy = body[1:]
        return (pts[:, 0] < ux) & (pts[:, 0] > lx) & (pts[:, 1] < uy) & (pts[:, 1] > ly)

    @staticmethod
    def _centre(body):
        if body[0] == "circle":
            return np.array(body[1:3])
        return (np.array(body[1:3]) + np.array(body[3:5])) / 2.

    @staticmethod
    def _box_edges(body):
        lx, ly, ux, uy = body[1:]
        A = [np.array(p) for p in ((lx, ly), (ux, ly), (ux, uy), (lx, uy))]
        return [(A[k], A[(k + 1) % 4]) for k in range(4)]

    def _nearest_intersect(self, body, p):
        if body[0] == "circle":
            c = np.array(body[1:3])
            d = p - c
            return d / np.sqrt(d[0] * d[0] + d[1] * d[1]) * body[3] + c
        best, bd = None, np.inf
        for a, b in self._box_edges(body):
            t, r = b - a, p - a
            dd = r[0] * t[0] + r[1] * t[1]
            ms = t[0] * t[0] + t[1] * t[1]
            if dd >= 0. and dd <= ms:  # LineSegment2D::isBounded
                rxc = (a + (dd * t) / ms) - p
            else:
                ea, eb = a - p, b - p
                rxc = (a if ea[0] * ea[0] + ea[1] * ea[1] < eb[0] * eb[0] + eb[1] * eb[1] else b) - p
            d2 = rxc[0] * rxc[0] + rxc[1] * rxc[1]
            if d2 < bd:
                best, bd = p + rxc, d2
        return best

    def _wall_normal(self, body, bp):
        if body[0] == "circle":
            w = np.array(body[1:3]) - bp
            return w / np.sqrt(w[0] * w[0] + w[1] * w[1])
        best, bd = None, np.inf
        for a, b in self._box_edges(body):
            t, r = b - a, bp - a
            dd = r[0] * t[0] + r[1] * t[1]
            ms = t[0] * t[0] + t[1] * t[1]
            if dd >= 0. and dd <= ms:
                q = r - (dd * t) / ms
                d2 = q[0] * q[0] + q[1] * q[1]
            else:
                ea, eb = a - bp, b - bp
                d2 = min(ea[0] * ea[0] + ea[1] * ea[1], eb[0] * eb[0] + eb[1] * eb[1])
            if d2 < bd:
                best, bd = (a, b), d2
        a, b = best
        t = b - a
        n = np.array([t[1], -t[0]])
        mid = (a + b) / 2.
        cen = self._centre(body) - mid
        u = n / np.sqrt(n[0] * n[0] + n[1] * n[1])
        return u if n[0] * cen[0] + n[1] * cen[1] > 0. else -u

    def _stencil(self, c, body):
        xc = self.cc[c]
        bp = self._nearest_intersect(body, xc)
        ip = 2. * bp - xc
        nw = self._wall_normal(body, bp)
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
        for o, body in enumerate(self.bodies):
            cx, cy = self._centre(body)
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

    def velocity_bcs(self, velocities, omegas=None):
        """GhostCellImmersedBoundary::velocityBcs: Vector2D rows in the block
        layout.  Body velocity at a point (Motion::velocity, Motion.h:31-33):
        vel + omega (pt - pos).tangentVec(), pos the body's centroid; the
        ghost rows take it at bp, the solid rows at the cell centroid."""
        N = self.m.n_cells
        e = O.Eqn.create(2 * N, 5)
        for c in range(N):
            if self.types[c] == FLUID:
                continue
            o = self.owner[c]
            v = velocities[o]
            w = 0.0 if omegas is None else float(omegas[o])
            pt = self.stencils[c]["bp"] if self.types[c] == GHOST else self.cc[c]
            d = pt - self._centre(self.bodies[o])
            u = (v[0] + (-d[1]) * w, v[1] + d[0] * w)
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
