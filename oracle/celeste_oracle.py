"""TEST INFRASTRUCTURE ONLY.  The CELESTE surface-tension model
(src/2D/Unstructured/FiniteVolume/Multiphase/Celeste.cpp, CelesteStencil.cpp,
SurfaceTensionForce.cpp, SurfaceTensionForceSmoothingKernel.cpp) restated on
the CPU in numpy, the checker for the device model in the multiphase driver.
Only tests/ may use it.

Per call of face_forces (Celeste::computeFaceInterfaceForces, Celeste.cpp:9-20):
  gammaTilde = A sum_k gamma_k K(x_k - x) V_k over the cells whose centroid
      lies strictly inside the circle of radius `kernel_width` (SmoothingKernel,
      SurfaceTensionForceSmoothingKernel.cpp:3-59; pow8: (eps^2 - r^2)^4,
      pow6: (eps^2 - r^2)^3, peskin: kcos(dx) kcos(dy)); A = 1 / sum K V;
      boundary faces take the cell value (zero gradient, the default type)
  gradGammaTilde = rows 3, 4 of pinv(A_s) b with the weighted stencil
      (CelesteStencil.cpp:5-55, 112-137): rows for the face neighbours, then
      the node (diagonal) neighbours, then -- for cells touching the boundary --
      the neighbours' boundary faces and the cell's own boundary faces;
      columns (dx^2/2, dy^2/2, dx dy, dx, dy) / s, s = |dx|^2 (weighted) or 1
  n = -gradGammaTilde / |gradGammaTilde| where |.|^2 >= eps^2 (eps = 1e-8), else
      0; boundary faces from the contact angle theta: n_f = ns cos(theta) +
      ts sin(theta) (SurfaceTensionForce.cpp:48-78)
  kappa = div(n) with the unweighted stencil where n != 0 on the cell, its
      face neighbours and its diagonal neighbours, else 0; faces: volume
      weighted or one-sided (Celeste.cpp:53-98)
  fst_f = sigma kappa_f gradGamma_f on every face

The least-squares pseudo-inverse is numpy's (SVD), where the reference calls
LAPACK dgels (QR, Math/Matrix.cpp:231-237): the same matrix, so the device
results agree to rounding, not bit for bit.  The kernel's cells are summed in
ascending cell id (the reference takes the order of its R-tree query).
Bit-level parity with the reference itself: parity unpinned (LAPACK and the
Boost R-tree are absent); what pins this restatement is in
tests/test_oracle_celeste.py (exactness for quadratic fields, kernel
normalisation, the curvature of a disc, contact-angle normals).
"""
import numpy as np

EPS = 1e-8  # SurfaceTensionForce.h: eps_ = 1e-8


class CelesteOracle:
    def __init__(self, mesh, cptr, cind, kernel_width, kernel_type="pow8"):
        self.m = mesh
        N, F = mesh.n_cells, mesh.n_faces
        f = mesh.faces()
        self.fl, self.fr = f["lcell"], f["rcell"]
        self.fc, self.sf = f["centroid"], f["sf"]
        self.vol, self.cc = mesh.cells()
        L = mesh.links()
        iptr, iface, inb, bptr, bface = L["iptr"], L["iface"], L["inb"], L["bptr"], L["bface"]
        self.iptr, self.iface, self.inb, self.bptr, self.bface = iptr, iface, inb, bptr, bface
        cptr, cind = np.asarray(cptr), np.asarray(cind)
        # Cell::diagonals (FiniteVolumeGrid2D.cpp:423-430): for each node of the
        # cell, each cell of that node (creation order) that is neither the cell
        # nor a face neighbour
        node_cells = [[] for _ in range(int(cind.max()) + 1)]
        for c in range(N):
            for k in range(cptr[c], cptr[c + 1]):
                node_cells[cind[k]].append(c)
        self.diag = []
        for c in range(N):
            nbs = set(inb[iptr[c]:iptr[c + 1]].tolist())
            d = []
            for k in range(cptr[c], cptr[c + 1]):
                for kc in node_cells[cind[k]]:
                    if kc != c and kc not in nbs:
                        d.append(kc)
            self.diag.append(d)
        # stencils (cells, faces) per cell, CelesteStencil.cpp:9-31
        self.st_cells, self.st_faces = [], []
        for c in range(N):
            cells, faces = [], []
            has_bd = bptr[c + 1] > bptr[c]
            for k in range(iptr[c], iptr[c + 1]):
                nb = inb[k]
                cells.append(nb)
                if has_bd:
                    faces.extend(bface[bptr[nb]:bptr[nb + 1]].tolist())
            cells.extend(self.diag[c])
            faces.extend(bface[bptr[c]:bptr[c + 1]].tolist())
            self.st_cells.append(np.array(cells, np.int64))
            self.st_faces.append(np.array(faces, np.int64))
        self.pinv_w = [self._pinv(c, True) for c in range(N)]
        self.pinv_u = [self._pinv(c, False) for c in range(N)]
        # smoothing kernels
        self.eps = float(kernel_width)
        self.kcells, self.kw, self.A = [], [], np.empty(N)
        for c in range(N):
            d = self.cc - self.cc[c]
            inside = np.where(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] < self.eps * self.eps)[0]
            w = np.array([self._kernel(d[j], kernel_type) for j in inside])
            self.kcells.append(inside)
            self.kw.append(w)
            a = 0.
            for j, wj in zip(inside, w):
                a += wj * self.vol[j]
            self.A[c] = 1. / a

    def _kernel(self, dx, kind):
        e = self.eps
        if kind == "peskin":
            kc = lambda x: e * (1. + np.cos(np.pi * x / e)) if x < e else 0.
            return kc(dx[0]) * kc(dx[1])
        r = np.sqrt(dx[0] * dx[0] + dx[1] * dx[1])
        p = 4 if kind == "pow8" else 3
        return (e ** 2 - r ** 2) ** p if r < e else 0.

    def _rows(self, c, weighted):
        x0 = self.cc[c]
        rows = []
        for k in self.st_cells[c]:
            r = self.cc[k] - x0
            s = r[0] * r[0] + r[1] * r[1] if weighted else 1.
            rows.append((r[0] * r[0] / 2. / s, r[1] * r[1] / 2. / s, r[0] * r[1] / s, r[0] / s, r[1] / s))
        for k in self.st_faces[c]:
            r = self.fc[k] - x0
            s = r[0] * r[0] + r[1] * r[1] if weighted else 1.
            rows.append((r[0] * r[0] / (2. * s), r[1] * r[1] / (2. * s), r[0] * r[1] / s, r[0] / s, r[1] / s))
        return np.array(rows)

    def _pinv(self, c, weighted):
        return np.linalg.pinv(self._rows(c, weighted))[3:5]

    def _b(self, c, phi_cell, phi_face, weighted):
        """(phi_k - phi_c) / s per stencil row (columns: components)."""
        x0 = self.cc[c]
        out = []
        for k in self.st_cells[c]:
            r = self.cc[k] - x0
            s = r[0] * r[0] + r[1] * r[1] if weighted else 1.
            out.append((phi_cell[k] - phi_cell[c]) / s)
        for k in self.st_faces[c]:
            r = self.fc[k] - x0
            s = r[0] * r[0] + r[1] * r[1] if weighted else 1.
            out.append((phi_face[k] - phi_cell[c]) / s)
        return np.array(out)

    def smooth(self, gamma):
        N, F = self.m.n_cells, self.m.n_faces
        gt = np.empty(N)
        for c in range(N):
            acc = 0.
            for j, w in zip(self.kcells[c], self.kw[c]):
                acc += gamma[j] * w * self.vol[j]
            gt[c] = self.A[c] * acc
        gtf = np.zeros(F)
        b = self.fr < 0
        gtf[b] = gt[self.fl[b]]
        return gt, gtf

    def grad(self, phi_cell, phi_face):
        """Weighted-stencil gradient (Celeste::Stencil::grad)."""
        N = self.m.n_cells
        g = np.empty((N, 2))
        for c in range(N):
            g[c] = self.pinv_w[c] @ self._b(c, phi_cell, phi_face, True)
        return g

    def div(self, u_cell, u_face):
        """Unweighted-stencil divergence (Celeste::Stencil::div / kappa)."""
        N = self.m.n_cells
        d = np.empty(N)
        for c in range(N):
            b = self._b(c, u_cell, u_face, False)
            q = self.pinv_u[c] @ b
            d[c] = q[0, 0] + q[1, 1]
        return d

    def normals(self, gt_grad, cos_t, sin_t):
        N, F = self.m.n_cells, self.m.n_faces
        m2 = gt_grad[:, 0] ** 2 + gt_grad[:, 1] ** 2
        nC = np.zeros((N, 2))
        ok = m2 >= EPS * EPS
        mag = np.sqrt(m2[ok])
        nC[ok] = -(gt_grad[ok] / mag[:, None])
        nF = np.zeros((F, 2))
        for f in np.where(self.fr < 0)[0]:
            l = self.fl[f]
            if nC[l, 0] ** 2 + nC[l, 1] ** 2 == 0.:
                continue
            sf = self.sf[f]
            d = self.fc[f] - self.cc[l]
            out = sf if d[0] * sf[0] + d[1] * sf[1] > 0 else -sf
            ns = -(out / np.sqrt(out[0] ** 2 + out[1] ** 2))
            t = (nC[l] @ ns) / (ns @ ns)
            tv = nC[l] - t * ns
            with np.errstate(invalid="ignore", divide="ignore"):
                ts = tv / np.sqrt(tv[0] ** 2 + tv[1] ** 2)
            nF[f] = nC[l] if np.any(np.isnan(ts)) else ns * cos_t[f] + ts * sin_t[f]
        return nC, nF

    def curvature(self, nC, nF, kappa_face_prev):
        N, F = self.m.n_cells, self.m.n_faces
        nz = (nC[:, 0] ** 2 + nC[:, 1] ** 2) != 0.
        d = self.div(nC, nF)
        kC = np.zeros(N)
        for c in range(N):
            nbs = self.inb[self.iptr[c]:self.iptr[c + 1]]
            if nz[c] and nz[nbs].all() and all(nz[k] for k in self.diag[c]):
                kC[c] = d[c]
        kF = np.array(kappa_face_prev, float).copy()
        for f in range(F):
            a, b = self.fl[f], self.fr[f]
            if b >= 0:
                if nz[a] and nz[b]:
                    g = self.vol[b] / (self.vol[b] + self.vol[a])
                    kF[f] = g * kC[a] + (1. - g) * kC[b]
                elif nz[a]:
                    kF[f] = kC[a]
                elif nz[b]:
                    kF[f] = kC[b]
                else:
                    kF[f] = 0.
            elif nz[a]:
                kF[f] = kC[a]
        return kC, kF

    def face_forces(self, gamma, grad_gamma_face, sigma, cos_t, sin_t, kappa_face_prev):
        gt, gtf = self.smooth(gamma)
        ggt = self.grad(gt, gtf)
        nC, nF = self.normals(ggt, cos_t, sin_t)
        kC, kF = self.curvature(nC, nF, kappa_face_prev)
        fst = (sigma * kF)[:, None] * np.asarray(grad_gamma_face).reshape(-1, 2)
        return dict(gamma_tilde=gt, grad_gamma_tilde=ggt, n_cell=nC, n_face=nF, kappa_cell=kC, kappa_face=kF,
                    fst_face=fst)
