"""TEST INFRASTRUCTURE ONLY.  SIMPLE outer iterations restated on the CPU with
the FP64 oracle's operators (phase_oracle.cpp) and direct solves (scipy), the
checker for the device SIMPLE driver (phase_amd/csrc/simple.hip).  Only tests/
may use it.

The algorithm is the README's SIMPLE equations (README.md:27-37) with the
implicit under-relaxation kept in the comment of
ScalarFiniteVolumeEquation.cpp:45-55 and collocated Rhie-Chow face velocities:

  p_f   = interpolateFaces(p) (zero gradient on the boundary); grad p by
          ScalarGradient (ScalarGradient.cpp:34-74)
  uEqn  = rho div(u_f, u) == laplacian(mu, u) - src(grad p); relax(alpha_u)
  d_P   = V_P / a_P (relaxed x-momentum diagonal); u* = solve(uEqn)
  u_f   = avg(u*) + avg(d) (avg(grad p) - grad_f p)       (interior faces)
  p'    = solve(laplacian(rho avg(d), p') == src::div(u_f)), p'_pin = 0
  u    -= d grad p';  u_f -= avg(d) grad_f p' (interior);  p += alpha_p p'

PISO (Issa 1986) uses the same pieces per time step: the momentum predictor
with fv::ddt(rho, u, dt) and no relaxation, then several pressure corrections
(alpha_p = 1), each from the latest velocity and pressure gradient.

Pinned by tests/test_oracle_fstep.py: SIMPLE and PISO marched to steady state
on the Re = 100 cavity land on the centreline data of Ghia, Ghia & Shin
(1982), and a PISO transient of the impulsively started cavity converges (first
order in dt) to the FractionalStep transient of the same flow, whose own
restatement is pinned by closed forms.  `plant` (tests only) plants one mistake
so that the tests can show each pin catches it: "flip_src" adds src(grad p)
instead of subtracting it, "ddt_half" halves the ddt coefficient.

Face averages use Face::distanceWeight (Face.cpp:60-64).  Elementwise
arithmetic is written in the device kernels' order, so only the linear solves
(direct here, BiCGStab there) differ.
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


def _diag_of(rp, ci, v, row):
    for j in range(rp[row], rp[row + 1]):
        if ci[j] == row:
            return v[j]
    return 0.0


class SimpleOracle:
    def __init__(self, mesh, rho, mu, alpha_u, alpha_p, u_bc, u_boundary, pin=0, plant=None):
        self.m = mesh
        self.rho, self.mu, self.au, self.ap, self.pin = rho, mu, alpha_u, alpha_p, pin
        self.plant = plant
        N, F = mesh.n_cells, mesh.n_faces
        f = mesh.faces()
        self.fl, self.fr = f["lcell"], f["rcell"]
        self.fc = f["centroid"].reshape(F, 2)
        self.vol, self.cc = mesh.cells()
        self.ubc = np.asarray(u_bc, np.int32)
        wall = np.asarray(u_boundary, float).reshape(F, 2).copy()
        wall[self.fr >= 0] = 0.0
        self.uwall = wall
        self.pbc = np.full(F, O.BC_NORMAL_GRADIENT, np.int32)
        self.uC = np.zeros((N, 2))
        self.pC = np.zeros(N)
        self.uF = wall.copy()
        inner = self.fr >= 0
        l, r = self.fl[inner], self.fr[inner]
        a = self.fc[inner] - self.cc[r]
        b = self.fc[inner] - self.cc[l]
        l1 = np.sqrt(a[:, 0] * a[:, 0] + a[:, 1] * a[:, 1])
        l2 = np.sqrt(b[:, 0] * b[:, 0] + b[:, 1] * b[:, 1])
        self.inner, self.il, self.ir = inner, l, r
        self.g = l1 / (l1 + l2)

    def iterate(self, n):
        """n SIMPLE outer iterations; rows of (max |imbalance|, max |p'|)."""
        hist = []
        for _ in range(n):
            self._pressure_gradient()
            self._momentum(0.0, self.au)
            hist.append(self._correct(self.ap))
        return np.array(hist)

    def piso(self, dt, n=1, correctors=2):
        """n PISO time steps: momentum predictor with ddt, no relaxation, then
        `correctors` pressure corrections from the latest u and grad p."""
        hist = []
        for _ in range(n):
            self.u0C = self.uC.copy()
            self._pressure_gradient()
            self._momentum(dt, 1.0)
            for k in range(correctors):
                if k > 0:
                    self._pressure_gradient()
                last = self._correct(1.0)
            hist.append(last)
        return np.array(hist)

    def _pressure_gradient(self):
        m, F = self.m, self.m.n_faces
        pF = O.interpolate_faces(m, self.pC, np.zeros(F), bc_type=self.pbc)
        self.gF, self.gC = O.grad(m, self.pC, pF)

    def _momentum(self, dt, alpha_u):
        m, N = self.m, self.m.n_cells
        gsrc = O.src_src(m, self.gC.reshape(-1), ncomp=2)
        E = O.div(m, self.uF.reshape(-1), self.uF.reshape(-1), theta=1.0, bc_type=self.ubc,
                  phi_face=self.uwall.reshape(-1), phi0_cell=self.uC.reshape(-1), phi0_face=self.uwall.reshape(-1),
                  ncomp=2)
        if dt > 0.0:
            T = O.ddt(m, 2.0 * dt if self.plant == "ddt_half" else dt, self.u0C.reshape(-1), ncomp=2)
            T += E
            E = T
        if self.rho != 1.0:
            E.scale(self.rho)
        L = O.laplacian(m, 1, ncomp=2, gamma=self.mu, bc_type=self.ubc, phi_face=self.uwall.reshape(-1))
        if self.plant == "flip_src":
            L += gsrc
        else:
            L -= gsrc
        E.eq(L)
        if alpha_u < 1.0:
            E.relax(alpha_u, self.uC.T.reshape(-1).copy())
        rp, ci, v, rhs = E.get()
        self.d = np.array([self.vol[c] / _diag_of(rp, ci, v, c) for c in range(N)])
        x = spla.spsolve(_csr(rp, ci, v).tocsc(), -rhs)
        self.uC = x.reshape(2, N).T.copy()

    def _correct(self, alpha_p):
        m, F = self.m, self.m.n_faces
        d, gF, gC = self.d, self.gF, self.gC
        g, l, r = self.g, self.il, self.ir
        dd = g * d[l] + (1.0 - g) * d[r]
        uF = self.uwall.copy()
        for q in range(2):
            ub = g * self.uC[l, q] + (1.0 - g) * self.uC[r, q]
            gb = g * gC[l, q] + (1.0 - g) * gC[r, q]
            uF[self.inner, q] = ub + dd * (gb - gF[self.inner, q])
        df = d[self.fl].copy()
        df[self.inner] = dd
        self.uF = uF
        mdot = O.src_div(m, uF.reshape(-1))
        P = O.laplacian(m, 3, gamma_face=self.rho * df, bc_type=self.pbc, phi_face=np.zeros(F))
        P -= mdot
        rp, ci, v, rhs = P.get()
        A = _csr(rp, ci, v).tolil()
        A[self.pin, :] = 0.0
        A[self.pin, self.pin] = 1.0
        rhs = rhs.copy()
        rhs[self.pin] = 0.0
        pp = spla.spsolve(A.tocsc(), -rhs)
        ppF = O.interpolate_faces(m, pp, np.zeros(F), bc_type=self.pbc)
        gppF, gppC = O.grad(m, pp, ppF)
        self.uC = self.uC - d[:, None] * gppC
        self.pC = self.pC + alpha_p * pp
        self.uF[self.inner] = self.uF[self.inner] - dd[:, None] * gppF[self.inner]
        return np.abs(mdot).max(), np.abs(pp).max()
