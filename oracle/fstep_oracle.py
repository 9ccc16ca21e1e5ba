"""TEST INFRASTRUCTURE ONLY.  FractionalStep::solve restated on the CPU with the
FP64 oracle's operators and direct solves, the checker for the device
fractional-step driver (phase_amd/csrc/simple.hip).  Only tests/ may use it.

One time step (src/2D/Unstructured/Solvers/FractionalStep.cpp:36-117):
  u0 = u                                               (savePreviousTimeStep)
  uEqn = ddt(u, dt) + div(u, u, 0) == laplacian(mu/rho, u, 0.5) - src(grad p)
  u = solve(uEqn); u += dt grad p; u_f = interpolateFaces(u)
  pEqn = laplacian(dt, p) == src::div(u_f); p = solve(pEqn)
  p_f = setBoundaryFaces(p); grad p (ScalarGradient)
  u -= dt grad p (cells); u_f -= dt grad_f p (every face)
With no FIXED pressure face the pressure matrix is singular and one cell is
pinned to 0, as on the device.

Pinned by tests/test_oracle_fstep.py: a pressure-driven Poiseuille channel
(steady profile G y (H - y) / 2 mu to second order, exactly linear pressure),
a viscous shear mode decaying at the Crank-Nicolson rate of the theta = 0.5
viscous term, and a closed box at rest staying at rest.  `plant` (tests only)
plants one mistake so that the tests can show each pin catches it:
"drop_restore" skips u += dt grad p, "flip_src" adds src(grad p) instead of
subtracting it, "theta1" makes the viscous term fully implicit,
"save_per_corrector" takes the old-time fields again before every pass of a
step with several pressure correctors (the step then advances one dt per pass).

With `correctors` = n > 1 everything after savePreviousTimeStep runs n times
per step: each later pass assembles the momentum equation from the same
old-time fields with the latest face velocities and pressure gradient and
projects again (the device driver's numPressureCorrections).

With an immersed boundary this is FractionalStepGCIB::solve
(src/2D/Unstructured/Solvers/FractionalStepGCIB.cpp:14-56): the fv:: terms
cover the fluid cells only (their other rows removed), the ghost-cell rows of
oracle/ghostcell_oracle.py are added (velocityBcs to the momentum equation,
bcs(p) to the pressure equation), the momentum source is src(grad p / rho),
u += dt / rho grad p and u -= dt grad p touch the fluid cells only, and
src::div and the divergence check cover the fluid cells.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O


def _csr(rp, ci, v):
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = ci >= 0
    return sp.csr_matrix((v[keep], (rows[keep], ci[keep])), shape=(n, n))


class FractionalStepOracle:
    def __init__(self, mesh, rho, mu, u_bc, u_face0, p_bc=None, p_face0=None, pin=0, immersed=None, plant=None):
        """immersed: None or (GhostCellOracle, body velocities, pressure bc type
        on the bodies, its reference value[, body angular velocities])."""
        self.m, self.rho, self.mu, self.pin = mesh, rho, mu, pin
        self.ib = immersed
        self.plant = plant
        N, F = mesh.n_cells, mesh.n_faces
        self.ubc = np.asarray(u_bc, np.int32)
        self.pbc = (np.full(F, O.BC_NORMAL_GRADIENT, np.int32) if p_bc is None else np.asarray(p_bc, np.int32))
        self.uC = np.zeros((N, 2))
        self.pC = np.zeros(N)
        self.gpC = np.zeros((N, 2))
        self.gpF = np.zeros((F, 2))
        self.pF = np.zeros(F) if p_face0 is None else np.asarray(p_face0, float).copy()
        uf = np.asarray(u_face0, float).reshape(F, 2)
        self.uF = O.interpolate_faces(mesh, self.uC.reshape(-1), uf.reshape(-1), bc_type=self.ubc,
                                      ncomp=2).reshape(F, 2)

    def _solve(self, eqn, pin=-1):
        rp, ci, v, rhs = eqn.get()
        A = _csr(rp, ci, v)
        b = -rhs
        if pin >= 0:
            A = A.tolil()
            A[pin, :] = 0.0
            A[pin, pin] = 1.0
            b = b.copy()
            b[pin] = 0.0
        return spla.spsolve(A.tocsc(), b)

    def step(self, dt, correctors=1):
        """One time step; `correctors` passes of momentum, pressure and
        velocity correction from the same old-time fields (the device driver's
        numPressureCorrections)."""
        u0C, u0F = self.uC.copy(), self.uF.copy()
        for _ in range(correctors):
            if self.plant == "save_per_corrector":
                u0C, u0F = self.uC.copy(), self.uF.copy()
            div = self._pass(dt, u0C, u0F)
        return div

    def _pass(self, dt, u0C, u0F):
        m, N, F = self.m, self.m.n_cells, self.m.n_faces
        ib = self.ib
        if ib is not None:
            g, vel, pbc_ib, pref = ib[:4]
            omegas = ib[4] if len(ib) > 4 else None
            other = (g.types != 0).astype(np.int32)
            fluid = g.types == 0
        E = O.ddt(m, dt, u0C.reshape(-1), ncomp=2)
        E += O.div(m, self.uF.reshape(-1), u0F.reshape(-1), theta=0.0, bc_type=self.ubc,
                   phi_face=self.uF.reshape(-1), phi0_cell=u0C.reshape(-1), phi0_face=u0F.reshape(-1), ncomp=2)
        if ib is not None:
            E.remove_cells(N, other)
            E += g.velocity_bcs(vel, omegas=omegas)
        theta = 1.0 if self.plant == "theta1" else 0.5
        L = O.laplacian(m, 0, ncomp=2, gamma=self.mu / self.rho, theta=theta, bc_type=self.ubc,
                        phi_face=self.uF.reshape(-1), phi0_cell=u0C.reshape(-1), phi0_face=u0F.reshape(-1))
        if ib is None:
            if self.plant == "flip_src":
                L += O.src_src(m, self.gpC.reshape(-1), ncomp=2)
            else:
                L -= O.src_src(m, self.gpC.reshape(-1), ncomp=2)
        else:
            L.remove_cells(N, other)
            src = O.src_src(m, (self.gpC / self.rho).reshape(-1), ncomp=2)
            src[np.r_[~fluid, ~fluid]] = 0.0
            L -= src
        E.eq(L)
        self.uC = self._solve(E).reshape(2, N).T.copy()
        if ib is None:
            if self.plant != "drop_restore":
                self.uC = self.uC + dt * self.gpC
        else:
            self.uC[fluid] = self.uC[fluid] + (dt / self.rho) * self.gpC[fluid]
        self.uF = O.interpolate_faces(m, self.uC.reshape(-1), self.uF.reshape(-1), bc_type=self.ubc,
                                      ncomp=2).reshape(F, 2)
        if ib is None:
            P = O.laplacian(m, 1, gamma=dt, bc_type=self.pbc, phi_face=self.pF)
            P -= O.src_div(m, self.uF.reshape(-1))
        else:
            P = O.laplacian(m, 1, gamma=dt / self.rho, bc_type=self.pbc, phi_face=self.pF)
            P.remove_cells(N, other)
            P += g.bcs(pbc_ib, pref)
            P -= O.src_div(m, self.uF.reshape(-1)) * fluid
        self.pC = self._solve(P, self.pin)
        self.pF = O.interpolate_faces(m, self.pC, self.pF, bc_type=self.pbc)
        gF, gC = O.grad(m, self.pC, self.pF)
        self.gpF, self.gpC = gF, gC
        if ib is None:
            self.uC = self.uC + (-dt) * self.gpC
        else:
            self.uC[fluid] = self.uC[fluid] + (-dt) * self.gpC[fluid]
        self.uF = self.uF + (-dt) * self.gpF
        div = np.abs(O.src_div(m, self.uF.reshape(-1)))
        return div.max() if ib is None else div[fluid].max()
