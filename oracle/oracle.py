"""TEST INFRASTRUCTURE ONLY: ctypes front-end of the FP64 CPU oracle.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, as the checker and as the CPU baseline.  The product package
(phase_amd) never imports it.  The C++ restatement it loads
(oracle/phase_oracle.cpp) cites the reference file:line of every routine.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libphase_oracle.so")

BC_FIXED, BC_NORMAL_GRADIENT, BC_SYMMETRY, BC_OUTFLOW, BC_PARTIAL_SLIP = 0, 1, 2, 3, 4

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def use_library(path):
    """Load the restatement from `path` instead (the sanitizer build); call
    before the first oracle function."""
    global _LIB_PATH, _lib
    _LIB_PATH = path
    _lib = None


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    sig = {
        "orc_mesh_create": (_vp, [ctypes.c_int, _f64p, ctypes.c_int, _i32p, _i32p,
                                  ctypes.c_char_p, ctypes.c_int]),
        "orc_mesh_free": (None, [_vp]),
        "orc_mesh_sizes": (None, [_vp, ctypes.POINTER(ctypes.c_int)] + [ctypes.POINTER(ctypes.c_int)] * 2),
        "orc_mesh_faces": (None, [_vp, _i32p, _i32p, _i32p, _i32p, _f64p, _f64p]),
        "orc_mesh_cells": (None, [_vp, _f64p, _f64p]),
        "orc_mesh_links": (None, [_vp, _i32p, _i32p, _i32p, _i32p, _i32p]),
        "orc_eqn_create": (_vp, [ctypes.c_int, ctypes.c_int]),
        "orc_eqn_copy": (_vp, [_vp]),
        "orc_eqn_free": (None, [_vp]),
        "orc_eqn_rank": (ctypes.c_int, [_vp]),
        "orc_eqn_nslots": (ctypes.c_int, [_vp]),
        "orc_eqn_get": (None, [_vp, _i32p, _i32p, _f64p, _f64p]),
        "orc_eqn_add_coeff": (None, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
        "orc_eqn_set_coeff": (None, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
        "orc_eqn_add_rhs": (None, [_vp, ctypes.c_int, ctypes.c_double]),
        "orc_eqn_plus_eqn": (None, [_vp, _vp, ctypes.c_int]),
        "orc_eqn_minus_vector": (None, [_vp, _f64p]),
        "orc_eqn_eq_scalar": (None, [_vp, ctypes.c_double]),
        "orc_eqn_scale": (None, [_vp, ctypes.c_double]),
        "orc_eqn_relax": (None, [_vp, ctypes.c_double, _f64p]),
        "orc_eqn_scale_cells": (None, [_vp, ctypes.c_int, _f64p]),
        "orc_eqn_remove_cells": (None, [_vp, ctypes.c_int, _i32p]),
        "orc_laplacian": (_vp, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _vp, _vp,
                                ctypes.c_double, _vp, _vp, _vp, _vp, ctypes.c_char_p, ctypes.c_int]),
        "orc_ddt": (_vp, [_vp, ctypes.c_int, ctypes.c_double, _f64p]),
        "orc_div": (_vp, [_vp, ctypes.c_int, _f64p, _f64p, ctypes.c_double, _vp, _vp, _vp, _vp,
                          ctypes.c_char_p, ctypes.c_int]),
        "orc_src_div": (None, [_vp, _f64p, _f64p]),
        "orc_divc": (_vp, [_vp, ctypes.c_int, _f64p, _f64p, ctypes.c_double, _vp, _vp, _vp, _vp,
                           ctypes.c_char_p, ctypes.c_int]),
        "orc_dive": (_vp, [_vp, ctypes.c_int, _f64p, _f64p, ctypes.c_double, _vp, _vp, _vp, _vp, _vp,
                           ctypes.c_char_p, ctypes.c_int]),
        "orc_div_grad": (_vp, [_vp, _f64p, _f64p, _vp, _vp, ctypes.c_char_p, ctypes.c_int]),
        "orc_ddt_rho": (_vp, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp, _vp, _f64p, _vp]),
        "orc_face_to_cell": (None, [_vp, _f64p, _vp, _vp, _f64p]),
        "orc_cicsam_hc": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
        "orc_cicsam_uq": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
        "orc_cicsam_weights": (None, [_vp, _f64p, _f64p, _f64p, ctypes.c_double, _f64p]),
        "orc_cicsam_div": (_vp, [_vp, _f64p, _f64p, ctypes.c_double, _vp, _vp, _vp, _vp, ctypes.c_char_p,
                                 ctypes.c_int]),
        "orc_cicsam_momentum_flux": (None, [_vp, ctypes.c_double, ctypes.c_double, _f64p, _f64p, _f64p, _f64p,
                                            _f64p]),
        "orc_vof_properties": (None, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]),
        "orc_src_src": (None, [_vp, ctypes.c_int, _f64p, _f64p]),
        "orc_grad": (None, [_vp, _f64p, _f64p, _f64p, _f64p]),
        "orc_interpolate_faces": (None, [_vp, ctypes.c_int, _f64p, _vp, _f64p]),
        "orc_spmv": (None, [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]),
        "orc_ilu0": (ctypes.c_int, [ctypes.c_int, _i32p, _i32p, _f64p]),
        "orc_lu_solve": (None, [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p]),
        "orc_bicgstab_ilu0": (ctypes.c_int, [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p,
                                             ctypes.c_double, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_double)]),
        "orc_bicgstab_ilu0_omp": (ctypes.c_int, [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p,
                                                 ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_int),
                                                 ctypes.POINTER(ctypes.c_double)]),
        "orc_bicgstab_ilut": (ctypes.c_int, [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


class Mesh:
    """FiniteVolumeGrid2D::init(nodes, cptr, cind) restated on the CPU."""

    def __init__(self, xy, cptr, cind):
        xy = _f64(np.asarray(xy).reshape(-1))
        cptr, cind = _i32(cptr), _i32(cind)
        err = ctypes.create_string_buffer(256)
        self._h = lib().orc_mesh_create(len(xy) // 2, xy, len(cptr) - 1, cptr, cind, err, 256)
        if not self._h:
            raise RuntimeError(err.value.decode())
        self.n_nodes, self.n_cells = len(xy) // 2, len(cptr) - 1
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib().orc_mesh_sizes(self._h, a, b, c)
        self.n_faces, self.n_interior, self.n_boundary = a.value, b.value, c.value

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_mesh_free(self._h)
            self._h = None

    def faces(self):
        F = self.n_faces
        out = dict(lnode=np.empty(F, np.int32), rnode=np.empty(F, np.int32),
                   lcell=np.empty(F, np.int32), rcell=np.empty(F, np.int32),
                   centroid=np.empty(2 * F), sf=np.empty(2 * F))
        lib().orc_mesh_faces(self._h, out["lnode"], out["rnode"], out["lcell"], out["rcell"],
                             out["centroid"], out["sf"])
        out["centroid"] = out["centroid"].reshape(F, 2)
        out["sf"] = out["sf"].reshape(F, 2)
        return out

    def cells(self):
        N = self.n_cells
        vol, cen = np.empty(N), np.empty(2 * N)
        lib().orc_mesh_cells(self._h, vol, cen)
        return vol, cen.reshape(N, 2)

    def links(self):
        N = self.n_cells
        iptr = np.empty(N + 1, np.int32)
        bptr = np.empty(N + 1, np.int32)
        iface = np.empty(2 * self.n_interior, np.int32)
        inb = np.empty(2 * self.n_interior, np.int32)
        bface = np.empty(self.n_boundary, np.int32)
        lib().orc_mesh_links(self._h, iptr, iface, inb, bptr, bface)
        return dict(iptr=iptr, iface=iface, inb=inb, bptr=bptr, bface=bface)


class Eqn:
    """CrsEquation (padded CSR with -1 empty slots) restated on the CPU."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError("null equation")
        self._h = handle

    @classmethod
    def create(cls, n_rows, nnz_per_row=5):
        return cls(lib().orc_eqn_create(n_rows, nnz_per_row))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_eqn_free(self._h)
            self._h = None

    def copy(self):
        return Eqn(lib().orc_eqn_copy(self._h))

    @property
    def rank(self):
        return lib().orc_eqn_rank(self._h)

    def get(self):
        n, s = self.rank, lib().orc_eqn_nslots(self._h)
        rp, ci = np.empty(n + 1, np.int32), np.empty(s, np.int32)
        v, rhs = np.empty(s), np.empty(n)
        lib().orc_eqn_get(self._h, rp, ci, v, rhs)
        return rp, ci, v, rhs

    def add_coeff(self, row, col, val):
        lib().orc_eqn_add_coeff(self._h, row, col, val)

    def set_coeff(self, row, col, val):
        lib().orc_eqn_set_coeff(self._h, row, col, val)

    def add_rhs(self, row, val):
        lib().orc_eqn_add_rhs(self._h, row, val)

    def __iadd__(self, other):
        if isinstance(other, Eqn):
            lib().orc_eqn_plus_eqn(self._h, other._h, 1)
        else:
            lib().orc_eqn_minus_vector(self._h, -_f64(other))
        return self

    def __isub__(self, other):
        if isinstance(other, Eqn):
            lib().orc_eqn_plus_eqn(self._h, other._h, -1)
        else:
            lib().orc_eqn_minus_vector(self._h, _f64(other))
        return self

    def eq(self, other):
        """CrsEquation::operator== (scalar, vector or equation)."""
        if isinstance(other, Eqn):
            lib().orc_eqn_plus_eqn(self._h, other._h, -1)
        elif np.isscalar(other):
            lib().orc_eqn_eq_scalar(self._h, float(other))
        else:
            lib().orc_eqn_minus_vector(self._h, _f64(other))
        return self

    def scale(self, s):
        lib().orc_eqn_scale(self._h, float(s))
        return self

    def scale_cells(self, n_cells, field):
        """operator*(ScalarFiniteVolumeField, equation): each cell's rows times field(cell)."""
        lib().orc_eqn_scale_cells(self._h, int(n_cells), _f64(field))
        return self

    def remove_cells(self, n_cells, mask):
        lib().orc_eqn_remove_cells(self._h, int(n_cells), _i32(mask))
        return self

    def relax(self, alpha, phi_rows):
        lib().orc_eqn_relax(self._h, float(alpha), _f64(phi_rows))
        return self


def laplacian(mesh, variant, ncomp=1, gamma=1.0, gamma_face=None, gamma0_face=None, theta=1.0,
              bc_type=None, phi_face=None, phi0_cell=None, phi0_face=None):
    gamma_face, gamma0_face = _f64(gamma_face), _f64(gamma0_face)
    bc_type = _i32(bc_type)
    phi_face, phi0_cell, phi0_face = _f64(phi_face), _f64(phi0_cell), _f64(phi0_face)
    if variant in (2, 3) and gamma_face is None:
        raise ValueError("field-gamma variants need gamma_face")
    if variant == 2 and gamma0_face is None:
        gamma0_face = gamma_face
    if variant in (0, 2):
        if phi0_cell is None:
            phi0_cell = np.zeros(ncomp * mesh.n_cells)
        if phi0_face is None:
            phi0_face = np.zeros(ncomp * mesh.n_faces)
    if phi_face is None:
        phi_face = np.zeros(ncomp * mesh.n_faces)
    err = ctypes.create_string_buffer(256)
    h = lib().orc_laplacian(mesh._h, variant, ncomp, float(gamma), _ptr(gamma_face),
                            _ptr(gamma0_face), float(theta), _ptr(bc_type), _ptr(phi_face),
                            _ptr(phi0_cell), _ptr(phi0_face), err, 256)
    if not h:
        raise RuntimeError(err.value.decode())
    return Eqn(h)


def ddt(mesh, time_step, phi0_cell, ncomp=1):
    return Eqn(lib().orc_ddt(mesh._h, ncomp, float(time_step), _f64(phi0_cell)))


def div(mesh, u_face, u0_face=None, theta=1.0, bc_type=None, phi_face=None, phi0_cell=None,
        phi0_face=None, ncomp=1):
    u_face = _f64(u_face)
    u0_face = u_face if u0_face is None else _f64(u0_face)
    if phi_face is None:
        phi_face = np.zeros(ncomp * mesh.n_faces)
    if phi0_cell is None:
        phi0_cell = np.zeros(ncomp * mesh.n_cells)
    if phi0_face is None:
        phi0_face = np.zeros(ncomp * mesh.n_faces)
    bc_type, phi_face, phi0_cell, phi0_face = (_i32(bc_type), _f64(phi_face), _f64(phi0_cell),
                                               _f64(phi0_face))
    err = ctypes.create_string_buffer(256)
    h = lib().orc_div(mesh._h, ncomp, u_face, u0_face, float(theta), _ptr(bc_type), _ptr(phi_face),
                      _ptr(phi0_cell), _ptr(phi0_face), err, 256)
    if not h:
        raise RuntimeError(err.value.decode())
    return Eqn(h)


def ddt_rho(mesh, time_step, phi0_cell, rho=1.0, rho_cell=None, rho0_cell=None, cell_mask=None, ncomp=1):
    """fv::ddt with a scalar or a field density, optionally over a cell group."""
    rho_cell, rho0_cell, cell_mask = _f64(rho_cell), _f64(rho0_cell), _i32(cell_mask)
    return Eqn(lib().orc_ddt_rho(mesh._h, ncomp, float(time_step), float(rho), _ptr(rho_cell), _ptr(rho0_cell),
                                 _f64(phi0_cell), _ptr(cell_mask)))


def _eqn_or_raise(h, err):
    if not h:
        raise RuntimeError(err.value.decode())
    return Eqn(h)


def divc(mesh, u_face, u0_face=None, theta=1.0, bc_type=None, phi_face=None, phi0_cell=None, phi0_face=None,
         ncomp=1):
    """fv::divc (central)."""
    u_face = _f64(u_face)
    u0_face = u_face if u0_face is None else _f64(u0_face)
    z = lambda n: np.zeros(ncomp * n)
    phi_face = z(mesh.n_faces) if phi_face is None else phi_face
    phi0_cell = z(mesh.n_cells) if phi0_cell is None else phi0_cell
    phi0_face = z(mesh.n_faces) if phi0_face is None else phi0_face
    bc_type, phi_face, phi0_cell, phi0_face = (_i32(bc_type), _f64(phi_face), _f64(phi0_cell), _f64(phi0_face))
    err = ctypes.create_string_buffer(256)
    return _eqn_or_raise(lib().orc_divc(mesh._h, ncomp, u_face, u0_face, float(theta), _ptr(bc_type),
                                        _ptr(phi_face), _ptr(phi0_cell), _ptr(phi0_face), err, 256), err)


def dive(mesh, u0_face, u1_face, theta, phi0_cell, phi1_cell, phi0_face, phi1_face, bc_type=None, ncomp=1):
    """fv::dive (explicit: sources only)."""
    a = [_f64(x) for x in (phi0_cell, phi1_cell, phi0_face, phi1_face)]
    bc_type = _i32(bc_type)
    err = ctypes.create_string_buffer(256)
    return _eqn_or_raise(lib().orc_dive(mesh._h, ncomp, _f64(u0_face), _f64(u1_face), float(theta),
                                        _ptr(bc_type), *[_ptr(x) for x in a], err, 256), err)


def div_grad(mesh, phiu_face, grad_u, bc_type=None, u_face=None):
    """fv::div(phiU, gradU, u) (Vector2D; grad_u: xx, xy, yx, yy per cell)."""
    u_face = np.zeros(2 * mesh.n_faces) if u_face is None else u_face
    bc_type, u_face = _i32(bc_type), _f64(u_face)
    err = ctypes.create_string_buffer(256)
    return _eqn_or_raise(lib().orc_div_grad(mesh._h, _f64(phiu_face), _f64(grad_u), _ptr(bc_type),
                                            _ptr(u_face), err, 256), err)


def face_to_cell(mesh, face_vec, cell_weight=None, face_weight=None):
    """VectorFiniteVolumeField::faceToCell, plain or weighted; (N, 2)."""
    if (cell_weight is None) != (face_weight is None):
        raise ValueError("faceToCell takes both weights or neither")
    out = np.empty(2 * mesh.n_cells)
    cw, fw = _f64(cell_weight), _f64(face_weight)
    lib().orc_face_to_cell(mesh._h, _f64(face_vec), _ptr(cw), _ptr(fw), out)
    return out.reshape(-1, 2)


def cicsam_hc(g, co):
    return lib().orc_cicsam_hc(float(g), float(co))


def cicsam_uq(g, co):
    return lib().orc_cicsam_uq(float(g), float(co))


def cicsam_weights(mesh, u_face, gamma_cell, grad_cell, time_step):
    """cicsam::faceInterpolationWeights: beta per face (0 on boundary faces)."""
    beta = np.empty(mesh.n_faces)
    lib().orc_cicsam_weights(mesh._h, _f64(u_face), _f64(gamma_cell), _f64(grad_cell), float(time_step), beta)
    return beta


def cicsam_div(mesh, u_face, beta, theta, bc_type=None, gamma_face=None, gamma0_cell=None, gamma0_face=None):
    """cicsam::div(u, gamma, beta, theta) as an equation."""
    if gamma_face is None:
        gamma_face = np.zeros(mesh.n_faces)
    if gamma0_cell is None:
        gamma0_cell = np.zeros(mesh.n_cells)
    if gamma0_face is None:
        gamma0_face = np.zeros(mesh.n_faces)
    bc_type, gamma_face, gamma0_cell, gamma0_face = (_i32(bc_type), _f64(gamma_face), _f64(gamma0_cell),
                                                     _f64(gamma0_face))
    err = ctypes.create_string_buffer(256)
    h = lib().orc_cicsam_div(mesh._h, _f64(u_face), _f64(beta), float(theta), _ptr(bc_type), _ptr(gamma_face),
                             _ptr(gamma0_cell), _ptr(gamma0_face), err, 256)
    if not h:
        raise RuntimeError(err.value.decode())
    return Eqn(h)


def cicsam_momentum_flux(mesh, rho1, rho2, u_face, gamma_cell, gamma_face, beta):
    out = np.empty(2 * mesh.n_faces)
    lib().orc_cicsam_momentum_flux(mesh._h, float(rho1), float(rho2), _f64(u_face), _f64(gamma_cell),
                                   _f64(gamma_face), _f64(beta), out)
    return out


def vof_properties(mesh, rho1, rho2, mu1, mu2, gamma_cell, gamma_face):
    """(rho_cell, rho_face, mu_cell, mu_face) of updateProperties."""
    rc, rf = np.empty(mesh.n_cells), np.empty(mesh.n_faces)
    mc, mf = np.empty(mesh.n_cells), np.empty(mesh.n_faces)
    lib().orc_vof_properties(mesh._h, float(rho1), float(rho2), float(mu1), float(mu2), _f64(gamma_cell),
                             _f64(gamma_face), rc, rf, mc, mf)
    return rc, rf, mc, mf


def src_div(mesh, u_face):
    out = np.empty(mesh.n_cells)
    lib().orc_src_div(mesh._h, _f64(u_face), out)
    return out


def src_src(mesh, phi_cell, ncomp=1):
    out = np.empty(ncomp * mesh.n_cells)
    lib().orc_src_src(mesh._h, ncomp, _f64(phi_cell), out)
    return out


def grad(mesh, phi_cell, phi_face):
    gf, gc = np.empty(2 * mesh.n_faces), np.empty(2 * mesh.n_cells)
    lib().orc_grad(mesh._h, _f64(phi_cell), _f64(phi_face), gf, gc)
    return gf.reshape(-1, 2), gc.reshape(-1, 2)


def interpolate_faces(mesh, phi_cell, phi_face, bc_type=None, ncomp=1):
    out = _f64(phi_face).copy()
    bc_type = _i32(bc_type)
    lib().orc_interpolate_faces(mesh._h, ncomp, _f64(phi_cell), _ptr(bc_type), out)
    return out


def spmv(row_ptr, col_ind, vals, x):
    n = len(row_ptr) - 1
    y = np.empty(n)
    lib().orc_spmv(n, _i32(row_ptr), _i32(col_ind), _f64(vals), _f64(x), y)
    return y


def ilu0(row_ptr, col_ind, vals):
    """ILU(0) in place on a compact, column-sorted CSR; returns the factor values."""
    v = _f64(vals).copy()
    rc = lib().orc_ilu0(len(row_ptr) - 1, _i32(row_ptr), _i32(col_ind), v)
    if rc != 0:
        raise RuntimeError("zero or missing pivot in row %d" % (-rc - 1))
    return v


def lu_solve(row_ptr, col_ind, vals, b):
    """Forward/backward substitution with a sorted-CSR ILU factor: returns (y, x)."""
    n = len(row_ptr) - 1
    y, x = np.empty(n), np.empty(n)
    lib().orc_lu_solve(n, _i32(row_ptr), _i32(col_ind), _f64(vals), _f64(b), y, x)
    return y, x


def bicgstab_ilu0(row_ptr, col_ind, vals, b, x0=None, tol=1e-8, max_iters=5000):
    n = len(row_ptr) - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    it, rr = ctypes.c_int(), ctypes.c_double()
    rc = lib().orc_bicgstab_ilu0(n, _i32(row_ptr), _i32(col_ind), _f64(vals), _f64(b), x,
                                 float(tol), int(max_iters), ctypes.byref(it), ctypes.byref(rr))
    return x, it.value, rr.value, rc


def bicgstab_ilu0_omp(row_ptr, col_ind, vals, b, x0=None, tol=1e-8, max_iters=5000, threads=1):
    """bicgstab_ilu0 with OpenMP SpMV / dots / updates (CPU timing baseline)."""
    n = len(row_ptr) - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    it, rr = ctypes.c_int(), ctypes.c_double()
    rc = lib().orc_bicgstab_ilu0_omp(n, _i32(row_ptr), _i32(col_ind), _f64(vals), _f64(b), x, float(tol),
                                     int(max_iters), int(threads), ctypes.byref(it), ctypes.byref(rr))
    return x, it.value, rr.value, rc


def bicgstab_ilut(row_ptr, col_ind, vals, b, x0=None, tol=1e-8, max_iters=5000, drop_tol=1e-12,
                  fill_factor=10):
    n = len(row_ptr) - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    it, rr = ctypes.c_int(), ctypes.c_double()
    rc = lib().orc_bicgstab_ilut(n, _i32(row_ptr), _i32(col_ind), _f64(vals), _f64(b), x,
                                 float(tol), int(max_iters), float(drop_tol), int(fill_factor),
                                 ctypes.byref(it), ctypes.byref(rr))
    return x, it.value, rr.value, rc


def to_dense(row_ptr, col_ind, vals, n_cols=None):
    n = len(row_ptr) - 1
    A = np.zeros((n, n if n_cols is None else n_cols))
    for i in range(n):
        for j in range(row_ptr[i], row_ptr[i + 1]):
            if col_ind[j] >= 0:
                A[i, col_ind[j]] += vals[j]
    return A
