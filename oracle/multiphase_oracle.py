"""TEST INFRASTRUCTURE ONLY.  FractionalStepMultiphase::solve restated on the
CPU with the FP64 oracle's operators (phase_oracle.cpp) and direct solves
(scipy), the checker for the device multiphase driver (phase_amd/csrc/simple.hip,
phase_hip_vof_*).  Only tests/ may use it.

One time step (FractionalStepMultiphase.cpp:58-72); the surface-tension
force fst is the CELESTE model of celeste_oracle.py when `surface_tension`
is given, else 0:

  solveGammaEqn (:76-100): beta = cicsam weights(u_f, gamma, grad gamma, dt);
      gamma0 = gamma; fv::ddt(gamma, dt) + cicsam::div(u, gamma, beta, 0.5) == 0;
      interpolateFaces(gamma); grad gamma
  updateProperties (:175-218): rho0 = rho, mu0 = mu; rho, mu from gamma (cells,
      faces); grad rho on the faces; sg_f = dot(g, -x_f) grad rho_f;
      sg = faceToCell(rho, rho) of sg_f; fst_f = sigma kappa_f grad gamma_f,
      fst = faceToCell(rho, rho) of fst_f
  solveUEqn (:102-150): u1 = u0, u0 = u; gradP = faceToCell(rho, rho0) of
      gradP_f; rho * ddt(u, dt) + rho * dive(u, u, 0.5) ==
      laplacian(mu, u, 0.5) + src(fst + sg - gradP); u += dt / rho gradP;
      faces: volume-weighted u - dt / rho (fst + sg) plus dt / rho_f
      (fst_f + sg_f) (interior),
      by boundary type on the patches
  solvePEqn (:152-163): laplacian(dt / rho, p) == src::div(u); boundary faces;
      gradP_f; gradP = faceToCell(rho, rho) of gradP_f
  correctVelocity (:165-173): u -= dt / rho gradP (cells and faces)
With `correctors` = n > 1, solveUEqn (after the one shift of the old-time
fields), solvePEqn and correctVelocity run n times per step, as the device
driver's numPressureCorrections does.

Elementwise arithmetic is written in the device kernels' order, so only the
linear solves (direct here, BiCGStab there) differ.  Each operator used here
is pinned on its own (tests/test_oracle_*.py); the composition is pinned by
the invariants of tests/test_oracle_multiphase.py (divergence-free faces,
conserved volume, a fluid at rest, the Laplace pressure jump of a drop).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O


def _solve(e):
    rp, ci, v, rhs = e.get()
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = ci >= 0
    A = sp.csr_matrix((v[keep], (rows[keep], ci[keep])), shape=(n, n))
    return spla.spsolve(A.tocsc(), -rhs)


class MultiphaseOracle:
    def __init__(self, mesh, rho1, rho2, mu1, mu2, gravity, u_bc, u_boundary, p_bc, p_boundary, gamma_cell,
                 gamma_bc=None, surface_tension=None):
        """surface_tension: None or (celeste_oracle.CelesteOracle, sigma,
        contact angle per face in radians or None)."""
        m = self.m = mesh
        N, F = mesh.n_cells, mesh.n_faces
        self.rho1, self.rho2, self.mu1, self.mu2 = rho1, rho2, mu1, mu2
        self.g = np.asarray(gravity, float)
        f = mesh.faces()
        self.fl, self.fr = f["lcell"], f["rcell"]
        self.fc = f["centroid"].reshape(F, 2)
        self.sf = f["sf"].reshape(F, 2)
        self.vol, self.cc = mesh.cells()
        self.ubc = np.asarray(u_bc, np.int32)
        self.pbc = np.asarray(p_bc, np.int32)
        self.gbc = (np.full(F, O.BC_NORMAL_GRADIENT, np.int32) if gamma_bc is None
                    else np.asarray(gamma_bc, np.int32))
        self.uC = np.zeros((N, 2))
        self.uF = np.asarray(u_boundary, float).reshape(F, 2).copy()
        self.uF = O.interpolate_faces(m, self.uC.reshape(-1), self.uF.reshape(-1), bc_type=self.ubc,
                                      ncomp=2).reshape(F, 2)
        self.u0C, self.u1C = self.uC.copy(), self.uC.copy()
        self.u0F, self.u1F = self.uF.copy(), self.uF.copy()
        self.pC = np.zeros(N)
        self.pF = O.interpolate_faces(m, self.pC, np.asarray(p_boundary, float), bc_type=self.pbc)
        self.gpF = np.zeros((F, 2))
        self.gpC = np.zeros((N, 2))
        self.gC = np.asarray(gamma_cell, float).copy()
        self.gF = O.interpolate_faces(m, self.gC, np.zeros(F), bc_type=self.gbc)
        _, self.ggC = O.grad(m, self.gC, self.gF)
        self.rhoC = np.zeros(N)
        self.rhoF = np.zeros(F)
        self.muC = np.zeros(N)
        self.muF = np.zeros(F)
        self.st = surface_tension
        self.fstF, self.fstC = np.zeros((F, 2)), np.zeros((N, 2))
        self.kappaF = np.zeros(F)
        self.st_fields = None
        if surface_tension is not None:
            ca = surface_tension[2]
            ca = np.full(F, np.pi / 2) if ca is None else np.asarray(ca, float)
            self.cos_t, self.sin_t = np.cos(ca), np.sin(ca)
        self._properties()

    def _properties(self):
        m = self.m
        self.rho0C, self.rho0F, self.mu0F = self.rhoC.copy(), self.rhoF.copy(), self.muF.copy()
        self.rhoC, self.rhoF, self.muC, self.muF = O.vof_properties(m, self.rho1, self.rho2, self.mu1, self.mu2,
                                                                    self.gC, self.gF)
        grF, _ = O.grad(m, self.rhoC, self.rhoF)
        # sg_f = dot(g, -x_f) grad rho_f
        h = self.g[0] * -self.fc[:, 0] + self.g[1] * -self.fc[:, 1]
        self.sgF = h[:, None] * grF
        self.sgC = O.face_to_cell(m, self.sgF.reshape(-1), self.rhoC, self.rhoF)
        if self.st is not None:
            model, sigma = self.st[0], self.st[1]
            ggF, _ = O.grad(m, self.gC, self.gF)
            r = model.face_forces(self.gC, ggF, sigma, self.cos_t, self.sin_t, self.kappaF)
            self.kappaF = r["kappa_face"]
            self.fstF = r["fst_face"]
            self.fstC = O.face_to_cell(m, self.fstF.reshape(-1), self.rhoC, self.rhoF)
            self.st_fields = r

    def step(self, dt, correctors=1):
        """One time step; `correctors` passes of solveUEqn (after the one
        shift of the old-time fields), solvePEqn and correctVelocity."""
        m = self.m
        N, F = m.n_cells, m.n_faces
        # solveGammaEqn
        beta = O.cicsam_weights(m, self.uF.reshape(-1), self.gC, self.ggC.reshape(-1), dt)
        g0C, g0F = self.gC.copy(), self.gF.copy()
        e = O.ddt(m, dt, g0C)
        e += O.cicsam_div(m, self.uF.reshape(-1), beta, 0.5, bc_type=self.gbc, gamma_face=g0F, gamma0_cell=g0C,
                          gamma0_face=g0F)
        self.gC = _solve(e)
        self.gF = O.interpolate_faces(m, self.gC, self.gF, bc_type=self.gbc)
        _, self.ggC = O.grad(m, self.gC, self.gF)
        self._properties()
        # solveUEqn
        self.u1C, self.u0C = self.u0C, self.uC.copy()
        self.u1F, self.u0F = self.u0F, self.uF.copy()
        for _ in range(correctors):
            div = self._pass(dt)
        return div

    def _pass(self, dt):
        m = self.m
        N = m.n_cells
        self.gpC = O.face_to_cell(m, self.gpF.reshape(-1), self.rhoC, self.rho0F)
        a = O.ddt(m, dt, self.u0C.reshape(-1), ncomp=2)
        a.scale_cells(N, self.rhoC)
        b = O.dive(m, self.u0F.reshape(-1), self.u1F.reshape(-1), 0.5, self.u0C.reshape(-1), self.u1C.reshape(-1),
                   self.u0F.reshape(-1), self.u1F.reshape(-1), bc_type=self.ubc, ncomp=2)
        b.scale_cells(N, self.rhoC)
        a += b
        lap = O.laplacian(m, 2, ncomp=2, gamma_face=self.muF, gamma0_face=self.mu0F, theta=0.5, bc_type=self.ubc,
                          phi_face=self.uF.reshape(-1), phi0_cell=self.u0C.reshape(-1),
                          phi0_face=self.u0F.reshape(-1))
        s = (self.fstC + self.sgC) - self.gpC
        src = O.src_src(m, s.reshape(-1), ncomp=2)
        lap -= -src  # + src::src(...): rhs += src
        a.eq(lap)
        x = _solve(a)
        self.uC = x.reshape(2, N).T.copy()
        k = dt / self.rhoC
        self.uC = self.uC + k[:, None] * self.gpC
        self.uF = self._face_velocity(dt)
        # solvePEqn
        p = O.laplacian(m, 3, gamma_face=dt / self.rhoF, bc_type=self.pbc, phi_face=self.pF)
        p -= O.src_div(m, self.uF.reshape(-1))
        self.pC = _solve(p)
        self.pF = O.interpolate_faces(m, self.pC, self.pF, bc_type=self.pbc)
        self.gpF, _ = O.grad(m, self.pC, self.pF)
        self.gpC = O.face_to_cell(m, self.gpF.reshape(-1), self.rhoC, self.rhoF)
        # correctVelocity
        self.uC = self.uC - (dt / self.rhoC)[:, None] * self.gpC
        self.uF = self.uF - (dt / self.rhoF)[:, None] * self.gpF
        return np.abs(O.src_div(m, self.uF.reshape(-1))).max()

    def _face_velocity(self, dt):
        l, r = self.fl, self.fr
        uF = self.uF.copy()
        sg, rc, rf, u = self.fstC + self.sgC, self.rhoC, self.rhoF, self.uC
        sgF = self.fstF + self.sgF
        for f in range(len(l)):
            a, b = l[f], r[f]
            kf = dt / rf[f]
            if b >= 0:
                g = self.vol[b] / (self.vol[b] + self.vol[a])  # Face::volumeWeight
                ka, kb = dt / rc[a], dt / rc[b]
                for q in range(2):
                    uF[f, q] = (g * (u[a, q] - ka * sg[a, q]) + (1. - g) * (u[b, q] - kb * sg[b, q])
                                + kf * sgF[f, q])
            elif self.ubc[f] == O.BC_NORMAL_GRADIENT:
                ka = dt / rc[a]
                for q in range(2):
                    uF[f, q] = u[a, q] - ka * sg[a, q] + kf * sgF[f, q]
            elif self.ubc[f] == O.BC_SYMMETRY:
                nx, ny = self.sf[f]
                t = (u[a, 0] * nx + u[a, 1] * ny) / (nx * nx + ny * ny)
                uF[f, 0] = u[a, 0] - t * nx
                uF[f, 1] = u[a, 1] - t * ny
        return uF
