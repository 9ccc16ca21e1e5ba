"""ctypes binding of libphase_hip.so (include/phase_hip.h)."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_lib", "libphase_hip.so")

# FiniteVolumeField<T>::BoundaryType (FiniteVolumeField.h:12)
BC_FIXED, BC_NORMAL_GRADIENT, BC_SYMMETRY, BC_OUTFLOW, BC_PARTIAL_SLIP = 0, 1, 2, 3, 4
MODE_FP64, MODE_MIXED, MODE_LEARNED = 0, 1, 2

_ERRORS = {-1: "EINVAL", -2: "EDEVICE", -3: "ESTATE", -4: "ENOCONV", -5: "EBREAKDOWN",
           -6: "EBOUNDARY"}


class PhaseHipError(RuntimeError):
    """The reference's Exception(class, method, description) (src/System/Exception.h)."""

    def __init__(self, code, method, msg):
        self.code = code
        super().__init__("HipSparseMatrixSolver::%s [%s]: %s" % (method, _ERRORS.get(code, code), msg))


_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)

_SIGS = {
    "phase_hip_last_error": (ctypes.c_char_p, [_vp]),
    "phase_hip_device_count": (ctypes.c_int, [_ip]),
    "phase_hip_solver_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int]),
    "phase_hip_solver_destroy": (ctypes.c_int, [_vp]),
    "phase_hip_solver_setup": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
    "phase_hip_solver_set_rank": (ctypes.c_int, [_vp, ctypes.c_int]),
    "phase_hip_solver_set": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _i32p, _f64p]),
    "phase_hip_solver_set_rhs": (ctypes.c_int, [_vp, _f64p]),
    "phase_hip_solver_set_guess": (ctypes.c_int, [_vp, _f64p]),
    "phase_hip_solver_solve": (ctypes.c_int, [_vp, _dp]),
    "phase_hip_solver_get_x": (ctypes.c_int, [_vp, _f64p]),
    "phase_hip_solver_n_iters": (ctypes.c_int, [_vp]),
    "phase_hip_solver_error": (ctypes.c_double, [_vp]),
    "phase_hip_solver_stats": (ctypes.c_int, [_vp, _ip, _ip, _ip, _dp]),
    "phase_hip_solver_export": (ctypes.c_int, [_vp, _ip, _vp, _vp, _vp, _vp]),
    "phase_hip_solver_spmv": (ctypes.c_int, [_vp, ctypes.c_int, _f64p, _f64p]),
    "phase_hip_solver_multigrid_info": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), _vp, ctypes.c_int]),
    "phase_hip_solver_multigrid_coarsest": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "phase_hip_solver_spmv_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, _f64p, _f64p]),
    "phase_hip_solver_ilu_export": (ctypes.c_int, [_vp, _ip, _vp, _vp, _vp, _vp]),
    "phase_hip_solver_time_kernels": (ctypes.c_int, [_vp, ctypes.c_int, _f64p, ctypes.c_int]),
    "phase_hip_solver_kernel_bytes": (ctypes.c_int, [_vp, _f64p, ctypes.c_int]),
    "phase_hip_solver_ordering_info": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "phase_hip_mesh_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _f64p,
                                             ctypes.c_int, _i32p, _i32p]),
    "phase_hip_mesh_destroy": (ctypes.c_int, [_vp]),
    "phase_hip_mesh_last_error": (ctypes.c_char_p, [_vp]),
    "phase_hip_mesh_sizes": (ctypes.c_int, [_vp, _ip, _ip, _ip, _ip]),
    "phase_hip_mesh_faces": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "phase_hip_mesh_geometry": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "phase_hip_assemble_laplacian": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_double, _vp, _vp,
                                                    ctypes.c_double, _vp, _vp, _vp, _vp]),
    "phase_hip_equation_sub_rhs": (ctypes.c_int, [_vp, _f64p]),
    "phase_hip_assemble_laplacian_device": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_double, _vp, _vp,
                                                           ctypes.c_double, _vp, _vp, _vp, _vp]),
    "phase_hip_equation_sub_rhs_device": (ctypes.c_int, [_vp, _vp]),
    "phase_hip_solver_zero_guess": (ctypes.c_int, [_vp]),
    "phase_hip_malloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "phase_hip_free": (ctypes.c_int, [_vp]),
    "phase_hip_memcpy": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t, ctypes.c_int]),
    "phase_hip_copy_device": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "phase_hip_equation_export_padded": (ctypes.c_int, [_vp, _ip, _vp, _vp, _vp, _vp]),
    "phase_hip_learned_capture": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "phase_hip_learned_n_pairs": (ctypes.c_int, [_vp, _ip]),
    "phase_hip_learned_pairs": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "phase_hip_learned_add_pairs": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "phase_hip_learned_train": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]),
    "phase_hip_learned_params": (ctypes.c_int, [_vp, _vp, _ip]),
    "phase_hip_learned_set_params": (ctypes.c_int, [_vp, _vp, ctypes.c_int]),
    "phase_hip_learned_eval": (ctypes.c_int, [_vp, _f64p, _f64p, _f64p, _f64p, _dp, _vp]),
    "phase_hip_learned_apply": (ctypes.c_int, [_vp, _f64p, _f64p]),
    "phase_hip_equation_create": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, ctypes.c_int]),
    "phase_hip_equation_destroy": (ctypes.c_int, [_vp]),
    "phase_hip_equation_rank": (ctypes.c_int, [_vp]),
    "phase_hip_equation_last_error": (ctypes.c_char_p, [_vp]),
    "phase_hip_eq_laplacian": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, _vp, _vp, ctypes.c_double, _vp,
                                              _vp, _vp, _vp]),
    "phase_hip_eq_ddt": (ctypes.c_int, [_vp, ctypes.c_double, _vp]),
    "phase_hip_eq_div": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, _vp, _vp, _vp, _vp]),
    "phase_hip_eq_cicsam_div": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, _vp, _vp, _vp, _vp]),
    "phase_hip_eq_ddt_rho": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _vp, _vp, _vp, _vp]),
    "phase_hip_eq_scale_cells": (ctypes.c_int, [_vp, _vp]),
    "phase_hip_eq_remove_cells": (ctypes.c_int, [_vp, _vp]),
    "phase_hip_eq_divc": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, _vp, _vp, _vp, _vp]),
    "phase_hip_eq_dive": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, _vp, _vp, _vp, _vp, _vp]),
    "phase_hip_eq_div_grad": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "phase_hip_cicsam_weights": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_double, _vp]),
    "phase_hip_field_face_to_cell": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "phase_hip_cicsam_momentum_flux": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _vp, _vp, _vp, _vp,
                                                      _vp]),
    "phase_hip_vof_properties": (ctypes.c_int, [_vp] + [ctypes.c_double] * 4 + [_vp] * 6),
    "phase_hip_eq_add": (ctypes.c_int, [_vp, _vp, ctypes.c_int]),
    "phase_hip_eq_scale": (ctypes.c_int, [_vp, ctypes.c_double]),
    "phase_hip_eq_sub_vector": (ctypes.c_int, [_vp, _vp]),
    "phase_hip_eq_eq_scalar": (ctypes.c_int, [_vp, ctypes.c_double]),
    "phase_hip_eq_relax": (ctypes.c_int, [_vp, ctypes.c_double, _vp]),
    "phase_hip_eq_copy": (ctypes.c_int, [_vp, _vp]),
    "phase_hip_eq_export": (ctypes.c_int, [_vp, _ip, _vp, _vp, _vp, _vp]),
    "phase_hip_solver_set_equation": (ctypes.c_int, [_vp, _vp]),
    "phase_hip_solver_get_x_device": (ctypes.c_int, [_vp, _vp]),
    "phase_hip_solver_set_guess_device": (ctypes.c_int, [_vp, _vp]),
    "phase_hip_field_src_div": (ctypes.c_int, [_vp, _vp, _vp]),
    "phase_hip_field_src_src": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "phase_hip_field_grad": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "phase_hip_field_interpolate_faces": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp]),
    "phase_hip_mesh_synchronize": (ctypes.c_int, [_vp]),
    "phase_hip_fstep_create": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, ctypes.c_double, ctypes.c_double, _i32p,
                                              _f64p, _vp, _vp, ctypes.c_int]),
    "phase_hip_fstep_destroy": (ctypes.c_int, [_vp]),
    "phase_hip_fstep_last_error": (ctypes.c_char_p, [_vp]),
    "phase_hip_fstep_setup": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
    "phase_hip_fstep_advance": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _vp]),
    "phase_hip_fstep_fields": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "phase_hip_fstep_immersed_boundary_motion": (ctypes.c_int, [_vp, _vp, _f64p, ctypes.c_int, ctypes.c_double]),
    "phase_hip_fstep_immersed_boundary": (ctypes.c_int, [_vp, _vp, _f64p, ctypes.c_int, ctypes.c_double]),
    "phase_hip_gcib_create": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, ctypes.c_int, _f64p]),
    "phase_hip_gcib_create_shapes": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, ctypes.c_int, _i32p, _f64p]),
    "phase_hip_gcib_destroy": (ctypes.c_int, [_vp]),
    "phase_hip_gcib_last_error": (ctypes.c_char_p, [_vp]),
    "phase_hip_gcib_cell_types": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "phase_hip_gcib_hydrodynamic_force": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _vp, _vp,
                                                          ctypes.c_double, ctypes.c_double, _vp]),
    "phase_hip_eq_gcib_bcs": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_double]),
    "phase_hip_eq_gcib_velocity_bcs": (ctypes.c_int, [_vp, _vp, _f64p]),
    "phase_hip_eq_gcib_motion_bcs": (ctypes.c_int, [_vp, _vp, _f64p]),
    "phase_hip_multiphase_create": (ctypes.c_int, [ctypes.POINTER(_vp), _vp] + [ctypes.c_double] * 6 +
                                    [_i32p, _f64p, _vp, _vp, _vp, _f64p, ctypes.c_int]),
    "phase_hip_multiphase_destroy": (ctypes.c_int, [_vp]),
    "phase_hip_multiphase_last_error": (ctypes.c_char_p, [_vp]),
    "phase_hip_multiphase_setup": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
    "phase_hip_multiphase_advance": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _vp]),
    "phase_hip_multiphase_fields": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "phase_hip_multiphase_surface_tension": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                                             _vp]),
    "phase_hip_multiphase_surface_fields": (ctypes.c_int, [_vp] + [_vp] * 7),
    "phase_hip_simple_create": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_double, _i32p, _f64p, ctypes.c_int]),
    "phase_hip_simple_destroy": (ctypes.c_int, [_vp]),
    "phase_hip_simple_last_error": (ctypes.c_char_p, [_vp]),
    "phase_hip_simple_setup": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
    "phase_hip_simple_iterate": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    "phase_hip_simple_fields": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "phase_hip_simple_piso_advance": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _vp]),
}

_lib = None


def library_path():
    return _LIB


def load_library():
    """Load libphase_hip.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        raise ImportError("libphase_hip.so is not built (%s); run __graft_entry__.build()" % _LIB)
    lib = ctypes.CDLL(_LIB)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def device_count():
    c = ctypes.c_int(0)
    load_library().phase_hip_device_count(ctypes.byref(c))
    return c.value


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


class DeviceArray:
    """A device allocation holding a copy of a host numpy array (HBM-resident
    operator input for the *_device entry points)."""

    def __init__(self, host, device=0):
        lib = load_library()
        self.host = np.ascontiguousarray(host)
        self.nbytes = self.host.nbytes
        self.ptr = _vp()
        rc = lib.phase_hip_malloc(device, self.nbytes, ctypes.byref(self.ptr))
        if rc != 0:
            raise PhaseHipError(rc, "malloc", "device allocation failed")
        rc = lib.phase_hip_memcpy(self.ptr, self.host.ctypes.data_as(_vp), self.nbytes, 1)
        if rc != 0:
            raise PhaseHipError(rc, "memcpy", "host to device copy failed")

    @classmethod
    def zeros(cls, n, dtype=np.float64, device=0):
        """A zero-filled device array of n elements."""
        return cls(np.zeros(n, dtype), device)

    def get(self):
        """Copy the device contents back into a new host array."""
        out = np.empty_like(self.host)
        rc = load_library().phase_hip_memcpy(out.ctypes.data_as(_vp), self.ptr, self.nbytes, 0)
        if rc != 0:
            raise PhaseHipError(rc, "memcpy", "device to host copy failed")
        return out

    def set(self, host):
        """Overwrite the device contents from a host array of the same size."""
        h = np.ascontiguousarray(host, dtype=self.host.dtype)
        if h.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        rc = load_library().phase_hip_memcpy(self.ptr, h.ctypes.data_as(_vp), self.nbytes, 1)
        if rc != 0:
            raise PhaseHipError(rc, "memcpy", "host to device copy failed")

    def copy_from(self, other):
        """Device-to-device copy of another DeviceArray of the same size."""
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        rc = load_library().phase_hip_copy_device(self.ptr, other.ptr, self.nbytes)
        if rc != 0:
            raise PhaseHipError(rc, "copy_device", "device to device copy failed")

    def free(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            load_library().phase_hip_free(self.ptr)
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _dp_or_none(a):
    return None if a is None else a.ptr


class HipMesh:
    """Unstructured 2D mesh held on the device (FiniteVolumeGrid2D::init)."""

    def __init__(self, xy, cptr, cind, device=0):
        lib = load_library()
        xy = _f64(np.asarray(xy).reshape(-1))
        cptr, cind = _i32(cptr), _i32(cind)
        h = _vp()
        rc = lib.phase_hip_mesh_create(ctypes.byref(h), device, len(xy) // 2, xy, len(cptr) - 1, cptr, cind)
        if rc != 0:
            raise PhaseHipError(rc, "meshCreate", "mesh creation failed")
        self._h = h
        a, b, c, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib.phase_hip_mesh_sizes(h, a, b, c, d)
        self.n_cells, self.n_faces, self.n_interior, self.n_boundary = a.value, b.value, c.value, d.value

    def close(self):
        if getattr(self, "_h", None):
            load_library().phase_hip_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def faces(self):
        F = self.n_faces
        out = {k: np.empty(F, np.int32) for k in ("lnode", "rnode", "lcell", "rcell")}
        load_library().phase_hip_mesh_faces(self._h, _p(out["lnode"]), _p(out["rnode"]), _p(out["lcell"]),
                                            _p(out["rcell"]))
        return out

    def geometry(self):
        N, F = self.n_cells, self.n_faces
        vol, cc, fc, sf = np.empty(N), np.empty(2 * N), np.empty(2 * F), np.empty(2 * F)
        rc = load_library().phase_hip_mesh_geometry(self._h, _p(vol), _p(cc), _p(fc), _p(sf))
        if rc != 0:
            raise PhaseHipError(rc, "geometry", load_library().phase_hip_mesh_last_error(self._h).decode())
        return vol, cc.reshape(N, 2), fc.reshape(F, 2), sf.reshape(F, 2)


    def _mcheck(self, rc, method):
        if rc != 0:
            raise PhaseHipError(rc, method, load_library().phase_hip_mesh_last_error(self._h).decode())

    def synchronize(self):
        self._mcheck(load_library().phase_hip_mesh_synchronize(self._h), "synchronize")

    # --- field kernels (device arrays in and out; Source.cpp, ScalarGradient.cpp,
    # FiniteVolumeField.tpp)
    def src_div(self, u_face, out):
        self._mcheck(load_library().phase_hip_field_src_div(self._h, u_face.ptr, out.ptr), "src::div")
        self.synchronize()

    def src_src(self, phi_cell, out, ncomp=1):
        self._mcheck(load_library().phase_hip_field_src_src(self._h, ncomp, phi_cell.ptr, out.ptr), "src::src")
        self.synchronize()

    def grad(self, phi_cell, phi_face, grad_face, grad_cell):
        self._mcheck(load_library().phase_hip_field_grad(self._h, phi_cell.ptr, phi_face.ptr, grad_face.ptr,
                                                         grad_cell.ptr), "grad")
        self.synchronize()

    def face_to_cell(self, face_vec, cell_vec, cell_weight=None, face_weight=None):
        self._mcheck(load_library().phase_hip_field_face_to_cell(self._h, face_vec.ptr, _dp_or_none(cell_weight),
                                                                 _dp_or_none(face_weight), cell_vec.ptr),
                     "faceToCell")
        self.synchronize()

    # --- volume of fluid (Cicsam.cpp, FractionalStepMultiphase.cpp)
    def cicsam_weights(self, u_face, gamma_cell, grad_gamma_cell, time_step, beta):
        self._mcheck(load_library().phase_hip_cicsam_weights(self._h, u_face.ptr, gamma_cell.ptr,
                                                             grad_gamma_cell.ptr, float(time_step), beta.ptr),
                     "cicsam::faceInterpolationWeights")
        self.synchronize()

    def cicsam_momentum_flux(self, rho1, rho2, u_face, gamma_cell, gamma_face, beta, rho_u):
        self._mcheck(load_library().phase_hip_cicsam_momentum_flux(self._h, float(rho1), float(rho2), u_face.ptr,
                                                                   gamma_cell.ptr, gamma_face.ptr, beta.ptr,
                                                                   rho_u.ptr), "cicsam::computeMomentumFlux")
        self.synchronize()

    def vof_properties(self, rho1, rho2, mu1, mu2, gamma_cell, gamma_face, rho_cell, rho_face, mu_cell, mu_face):
        self._mcheck(load_library().phase_hip_vof_properties(self._h, float(rho1), float(rho2), float(mu1),
                                                             float(mu2), gamma_cell.ptr, gamma_face.ptr,
                                                             rho_cell.ptr, rho_face.ptr, mu_cell.ptr, mu_face.ptr),
                     "updateProperties")
        self.synchronize()

    def interpolate_faces(self, phi_cell, phi_face, bc_type=None, ncomp=1):
        self._mcheck(load_library().phase_hip_field_interpolate_faces(self._h, ncomp, phi_cell.ptr,
                                                                      _dp_or_none(bc_type), phi_face.ptr),
                     "interpolateFaces")
        self.synchronize()


class HipEquation:
    """FiniteVolumeEquation<T> / CrsEquation held on the device; operator
    inputs are DeviceArray objects (HBM-resident)."""

    def __init__(self, mesh, ncomp=1):
        h = _vp()
        rc = load_library().phase_hip_equation_create(ctypes.byref(h), mesh._h, ncomp)
        if rc != 0:
            raise PhaseHipError(rc, "equationCreate", "equation creation failed")
        self._h = h
        self.mesh = mesh
        self.ncomp = ncomp
        self.rank = load_library().phase_hip_equation_rank(h)

    def _check(self, rc, method):
        if rc != 0:
            raise PhaseHipError(rc, method, load_library().phase_hip_equation_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            load_library().phase_hip_equation_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def laplacian(self, variant, gamma=1.0, gamma_face=None, gamma0_face=None, theta=1.0, bc_type=None,
                  phi_face=None, phi0_cell=None, phi0_face=None):
        d = _dp_or_none
        self._check(load_library().phase_hip_eq_laplacian(self._h, variant, float(gamma), d(gamma_face),
                                                          d(gamma0_face), float(theta), d(bc_type), d(phi_face),
                                                          d(phi0_cell), d(phi0_face)), "fv::laplacian")
        return self

    def ddt(self, time_step, phi0_cell):
        self._check(load_library().phase_hip_eq_ddt(self._h, float(time_step), phi0_cell.ptr), "fv::ddt")
        return self

    def div(self, u_face, u0_face, theta=1.0, bc_type=None, phi_face=None, phi0_cell=None, phi0_face=None):
        d = _dp_or_none
        self._check(load_library().phase_hip_eq_div(self._h, u_face.ptr, u0_face.ptr, float(theta), d(bc_type),
                                                    d(phi_face), d(phi0_cell), d(phi0_face)), "fv::div")
        return self

    def ddt_rho(self, time_step, phi0_cell, rho=1.0, rho_cell=None, rho0_cell=None, cell_mask=None):
        d = _dp_or_none
        self._check(load_library().phase_hip_eq_ddt_rho(self._h, float(time_step), float(rho), d(rho_cell),
                                                        d(rho0_cell), phi0_cell.ptr, d(cell_mask)), "fv::ddt")
        return self

    def divc(self, u_face, u0_face, theta=1.0, bc_type=None, phi_face=None, phi0_cell=None, phi0_face=None):
        d = _dp_or_none
        self._check(load_library().phase_hip_eq_divc(self._h, u_face.ptr, u0_face.ptr, float(theta), d(bc_type),
                                                     d(phi_face), d(phi0_cell), d(phi0_face)), "fv::divc")
        return self

    def dive(self, u0_face, u1_face, theta, phi0_cell, phi1_cell, phi0_face, phi1_face, bc_type=None):
        self._check(load_library().phase_hip_eq_dive(self._h, u0_face.ptr, u1_face.ptr, float(theta),
                                                     _dp_or_none(bc_type), phi0_cell.ptr, phi1_cell.ptr,
                                                     phi0_face.ptr, phi1_face.ptr), "fv::dive")
        return self

    def div_grad(self, phiu_face, grad_u, bc_type=None, u_face=None):
        self._check(load_library().phase_hip_eq_div_grad(self._h, phiu_face.ptr, grad_u.ptr, _dp_or_none(bc_type),
                                                         _dp_or_none(u_face)), "fv::div")
        return self

    def cicsam_div(self, u_face, beta, theta, bc_type=None, gamma_face=None, gamma0_cell=None, gamma0_face=None):
        d = _dp_or_none
        self._check(load_library().phase_hip_eq_cicsam_div(self._h, u_face.ptr, beta.ptr, float(theta), d(bc_type),
                                                           d(gamma_face), d(gamma0_cell), d(gamma0_face)),
                    "cicsam::div")
        return self

    def gcib_bcs(self, ib, bc_type, ref_value):
        """GhostCellImmersedBoundary::bcs(phi) (scalar equation)."""
        self._check(load_library().phase_hip_eq_gcib_bcs(self._h, ib._h, int(bc_type), float(ref_value)), "bcs")
        return self

    def gcib_velocity_bcs(self, ib, velocities, omegas=None):
        """GhostCellImmersedBoundary::velocityBcs(u): body velocities (and
        angular velocities about their centroids)."""
        v = np.asarray(velocities, float).reshape(-1, 2)
        if omegas is None:
            rc = load_library().phase_hip_eq_gcib_velocity_bcs(self._h, ib._h, _f64(v.reshape(-1)))
        else:
            m = _f64(np.c_[v, np.asarray(omegas, float)].reshape(-1))
            rc = load_library().phase_hip_eq_gcib_motion_bcs(self._h, ib._h, m)
        self._check(rc, "velocityBcs")
        return self

    def __iadd__(self, other):
        self._check(load_library().phase_hip_eq_add(self._h, other._h, 1), "operator+=")
        return self

    def __isub__(self, other):
        self._check(load_library().phase_hip_eq_add(self._h, other._h, -1), "operator-=")
        return self

    def scale_cells(self, field):
        """operator*(ScalarFiniteVolumeField, equation)."""
        self._check(load_library().phase_hip_eq_scale_cells(self._h, field.ptr), "operator*")
        return self

    def remove_cells(self, mask):
        self._check(load_library().phase_hip_eq_remove_cells(self._h, mask.ptr), "remove")
        return self

    def scale(self, s):
        self._check(load_library().phase_hip_eq_scale(self._h, float(s)), "operator*=")
        return self

    def sub_vector(self, v):
        self._check(load_library().phase_hip_eq_sub_vector(self._h, v.ptr), "operator-=")
        return self

    def eq_scalar(self, s):
        self._check(load_library().phase_hip_eq_eq_scalar(self._h, float(s)), "operator==")
        return self

    def relax(self, alpha, phi_rows):
        self._check(load_library().phase_hip_eq_relax(self._h, float(alpha), phi_rows.ptr), "relax")
        return self

    def copy_from(self, other):
        self._check(load_library().phase_hip_eq_copy(self._h, other._h), "copy")
        return self

    def export(self):
        """(rowPtr, colInd, vals, rhs) in the reference's CrsEquation layout."""
        n = ctypes.c_int()
        self._check(load_library().phase_hip_eq_export(self._h, n, None, None, None, None), "export")
        rp = np.empty(self.rank + 1, np.int32)
        ci, v, rhs = np.empty(n.value, np.int32), np.empty(n.value), np.empty(self.rank)
        self._check(load_library().phase_hip_eq_export(self._h, n, _p(rp), _p(ci), _p(v), _p(rhs)), "export")
        return rp, ci, v, rhs


class HipFractionalStep:
    """FractionalStep::solve (src/2D/Unstructured/Solvers/FractionalStep.cpp:36-117)
    on the device: one projection time step per call of advance()."""

    def __init__(self, mesh, rho, mu, u_bc, u_face0, p_bc=None, p_face0=None, pin_cell=0):
        h = _vp()
        self._pbc = None if p_bc is None else _i32(p_bc)
        self._pf0 = None if p_face0 is None else _f64(p_face0)
        rc = load_library().phase_hip_fstep_create(ctypes.byref(h), mesh._h, float(rho), float(mu), _i32(u_bc),
                                                   _f64(np.asarray(u_face0).reshape(-1)), _p(self._pbc),
                                                   _p(self._pf0), int(pin_cell))
        if rc != 0:
            raise PhaseHipError(rc, "fractionalStepCreate", "creation failed")
        self._h = h
        self.mesh = mesh

    def _check(self, rc, method):
        if rc != 0:
            raise PhaseHipError(rc, method, load_library().phase_hip_fstep_last_error(self._h).decode())

    def setup(self, key, value):
        self._check(load_library().phase_hip_fstep_setup(self._h, key.encode(), str(value).encode()), "setup")

    def immersed_boundary(self, ib, velocities, p_bc, p_ref=0.0, omegas=None):
        """FractionalStepGCIB: the ghost-cell rows of `ib` (HipGhostCellIB) in
        both equations from the next step on; `omegas` (one per body) makes
        the bodies rotate about their centroids."""
        self._ib = ib  # the driver keeps using its cell types and rows
        vel = np.asarray(velocities, float).reshape(-1, 2)
        if omegas is None:
            self._check(load_library().phase_hip_fstep_immersed_boundary(self._h, ib._h, _f64(vel.reshape(-1)),
                                                                         int(p_bc), float(p_ref)), "immersedBoundary")
        else:
            motion = _f64(np.c_[vel, np.asarray(omegas, float).reshape(-1)].reshape(-1))
            self._check(load_library().phase_hip_fstep_immersed_boundary_motion(self._h, ib._h, motion, int(p_bc),
                                                                                float(p_ref)), "immersedBoundary")

    def advance(self, dt, n=1):
        """n time steps; returns an (n, 3) history: momentum and pressure
        iterations, max divergence of the corrected face velocities."""
        hist = np.zeros((n, 3))
        self._check(load_library().phase_hip_fstep_advance(self._h, float(dt), int(n), _p(hist)), "advance")
        return hist

    def fields(self):
        N, F = self.mesh.n_cells, self.mesh.n_faces
        u, p, uf = np.empty(2 * N), np.empty(N), np.empty(2 * F)
        self._check(load_library().phase_hip_fstep_fields(self._h, _p(u), _p(p), _p(uf)), "fields")
        return u.reshape(N, 2), p, uf.reshape(F, 2)

    def close(self):
        if getattr(self, "_h", None):
            load_library().phase_hip_fstep_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass



class HipGhostCellIB:
    """GhostCellImmersedBoundary for circular objects ((cx, cy, r) each):
    cell types (0 fluid, 1 ghost, 2 solid) and the boundary rows."""

    def __init__(self, mesh, circles, boxes=()):
        """circles: (cx, cy, r) each; boxes: (lower x, lower y, upper x,
        upper y) each; bodies are the circles, then the boxes."""
        h = _vp()
        circles, boxes = list(circles), list(boxes)
        self._n_bodies = len(circles) + len(boxes)
        if boxes:
            kinds = _i32([0] * len(circles) + [1] * len(boxes))
            geom = _f64([list(map(float, c)) + [0.0] for c in circles] + [list(map(float, b)) for b in boxes])
            rc = load_library().phase_hip_gcib_create_shapes(ctypes.byref(h), mesh._h, self._n_bodies, kinds,
                                                             geom.reshape(-1))
        else:
            c = _f64(np.asarray(circles, float).reshape(-1))
            rc = load_library().phase_hip_gcib_create(ctypes.byref(h), mesh._h, len(c) // 3, c)
        if rc != 0:
            raise PhaseHipError(rc, "ghostCellCreate", "creation failed")
        self._h = h
        self.mesh = mesh

    def cell_types(self):
        N = self.mesh.n_cells
        t, o = np.empty(N, np.int32), np.empty(N, np.int32)
        rc = load_library().phase_hip_gcib_cell_types(self._h, _p(t), _p(o), 0)
        if rc != 0:
            raise PhaseHipError(rc, "cellTypes", "")
        return t, o

    def hydrodynamic_force(self, rho, mu, u_cell, p_cell, g=(0.0, 0.0)):
        """applyHydrodynamicForce: (n_bodies, 2) forces from device fields."""
        n = load_library()
        f = np.zeros((self._n_bodies, 2))
        rc = n.phase_hip_gcib_hydrodynamic_force(self._h, float(rho), float(mu), u_cell.ptr, p_cell.ptr,
                                                 float(g[0]), float(g[1]), _p(f))
        if rc != 0:
            raise PhaseHipError(rc, "hydrodynamicForce", n.phase_hip_gcib_last_error(self._h).decode())
        return f

    def cell_types_device(self, out):
        rc = load_library().phase_hip_gcib_cell_types(self._h, out.ptr, None, 1)
        if rc != 0:
            raise PhaseHipError(rc, "cellTypes", "")

    def close(self):
        if getattr(self, "_h", None):
            load_library().phase_hip_gcib_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipMultiphase:
    """FractionalStepMultiphase::solve (src/2D/Unstructured/Solvers/
    FractionalStepMultiphase.cpp:58-218) on the device, without the surface
    tension force: volume fraction (CICSAM), properties, momentum and pressure
    per call of advance().  gamma = 0 is phase 1 (rho1, mu1), 1 phase 2."""

    def __init__(self, mesh, rho1, rho2, mu1, mu2, gravity, u_bc, u_face0, gamma_cell, p_bc=None, p_face0=None,
                 gamma_bc=None, pin_cell=-1):
        h = _vp()
        self._pbc = None if p_bc is None else _i32(p_bc)
        self._pf0 = None if p_face0 is None else _f64(p_face0)
        self._gbc = None if gamma_bc is None else _i32(gamma_bc)
        gx, gy = (float(v) for v in gravity)
        rc = load_library().phase_hip_multiphase_create(
            ctypes.byref(h), mesh._h, float(rho1), float(rho2), float(mu1), float(mu2), gx, gy, _i32(u_bc),
            _f64(np.asarray(u_face0).reshape(-1)), _p(self._pbc), _p(self._pf0), _p(self._gbc), _f64(gamma_cell),
            int(pin_cell))
        if rc != 0:
            raise PhaseHipError(rc, "multiphaseCreate", "creation failed")
        self._h = h
        self.mesh = mesh

    def _check(self, rc, method):
        if rc != 0:
            raise PhaseHipError(rc, method, load_library().phase_hip_multiphase_last_error(self._h).decode())

    def setup(self, key, value):
        self._check(load_library().phase_hip_multiphase_setup(self._h, key.encode(), str(value).encode()), "setup")

    def advance(self, dt, n=1):
        """n time steps; returns an (n, 4) history: gamma, momentum and
        pressure iterations, max divergence of the corrected face velocities."""
        hist = np.zeros((n, 4))
        self._check(load_library().phase_hip_multiphase_advance(self._h, float(dt), int(n), _p(hist)), "advance")
        return hist

    KERNELS = {"peskin": 0, "pow6": 1, "pow8": 2}

    def surface_tension(self, sigma, kernel_radius, kernel_type="pow8", contact_angle=None):
        """Turn on the CELESTE surface-tension force (Celeste.cpp); contact
        angle per face in radians (None: pi/2)."""
        ca = None if contact_angle is None else _f64(contact_angle)
        self._check(load_library().phase_hip_multiphase_surface_tension(
            self._h, float(sigma), float(kernel_radius), self.KERNELS[kernel_type.lower()], _p(ca)),
            "surfaceTension")

    def surface_fields(self):
        """dict of the surface-tension model's last fields."""
        N, F = self.mesh.n_cells, self.mesh.n_faces
        out = dict(gamma_tilde=np.empty(N), n_cell=np.empty(2 * N), n_face=np.empty(2 * F), kappa_cell=np.empty(N),
                   kappa_face=np.empty(F), fst_face=np.empty(2 * F), fst_cell=np.empty(2 * N))
        keys = ("gamma_tilde", "n_cell", "n_face", "kappa_cell", "kappa_face", "fst_face", "fst_cell")
        self._check(load_library().phase_hip_multiphase_surface_fields(self._h, *[_p(out[k]) for k in keys]),
                    "surfaceFields")
        for k in ("n_cell", "n_face", "fst_face", "fst_cell"):
            out[k] = out[k].reshape(-1, 2)
        return out

    def fields(self):
        """(u cells (N, 2), p cells, u faces (F, 2), gamma cells)."""
        N, F = self.mesh.n_cells, self.mesh.n_faces
        u, p, uf, g = np.empty(2 * N), np.empty(N), np.empty(2 * F), np.empty(N)
        self._check(load_library().phase_hip_multiphase_fields(self._h, _p(u), _p(p), _p(uf), _p(g)), "fields")
        return u.reshape(N, 2), p, uf.reshape(F, 2), g

    def close(self):
        if getattr(self, "_h", None):
            load_library().phase_hip_multiphase_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

class HipSimple:
    """SIMPLE driver on the device (steady incompressible flow): momentum and
    pressure-correction equations of the README, Rhie-Chow face velocities."""

    def __init__(self, mesh, rho, mu, alpha_u, alpha_p, u_bc, u_boundary, pin_cell=0):
        h = _vp()
        rc = load_library().phase_hip_simple_create(ctypes.byref(h), mesh._h, float(rho), float(mu), float(alpha_u),
                                                    float(alpha_p), _i32(u_bc), _f64(np.asarray(u_boundary).reshape(-1)),
                                                    int(pin_cell))
        if rc != 0:
            raise PhaseHipError(rc, "simpleCreate", load_library().phase_hip_last_error(None).decode())
        self._h = h
        self.mesh = mesh

    def _check(self, rc, method):
        if rc != 0:
            raise PhaseHipError(rc, method, load_library().phase_hip_simple_last_error(self._h).decode())

    def setup(self, key, value):
        self._check(load_library().phase_hip_simple_setup(self._h, key.encode(), str(value).encode()), "setup")

    def iterate(self, n):
        """n outer iterations; returns an (n, 4) history: momentum and
        pressure-correction iterations, max |mass imbalance|, max |p'|."""
        hist = np.zeros((n, 4))
        self._check(load_library().phase_hip_simple_iterate(self._h, int(n), _p(hist)), "iterate")
        return hist

    def piso(self, dt, n=1, correctors=2):
        """n PISO time steps of size dt with `correctors` pressure corrections."""
        hist = np.zeros((n, 4))
        self._check(load_library().phase_hip_simple_piso_advance(self._h, float(dt), int(n), int(correctors),
                                                                 _p(hist)), "piso")
        return hist

    def fields(self):
        N, F = self.mesh.n_cells, self.mesh.n_faces
        u, p, uf = np.empty(2 * N), np.empty(N), np.empty(2 * F)
        self._check(load_library().phase_hip_simple_fields(self._h, _p(u), _p(p), _p(uf)), "fields")
        return u.reshape(N, 2), p, uf.reshape(F, 2)

    def close(self):
        if getattr(self, "_h", None):
            load_library().phase_hip_simple_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipSparseMatrixSolver:
    """SparseMatrixSolver plugin "hip" (src/Math/SparseMatrixSolver.h:11-62)."""

    def __init__(self, device=0, **setup):
        lib = load_library()
        h = _vp()
        rc = lib.phase_hip_solver_create(ctypes.byref(h), device)
        if rc != 0:
            raise PhaseHipError(rc, "create", lib.phase_hip_last_error(None).decode())
        self._h = h
        self.rank = 0
        for k, v in setup.items():
            self.setup(k, v)

    def _check(self, rc, method):
        if rc != 0:
            raise PhaseHipError(rc, method, load_library().phase_hip_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            load_library().phase_hip_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- SparseMatrixSolver interface
    def setup(self, key, value):
        self._check(load_library().phase_hip_solver_setup(self._h, key.encode(), str(value).encode()), "setup")

    def set_rank(self, rank):
        self._check(load_library().phase_hip_solver_set_rank(self._h, int(rank)), "setRank")
        self.rank = int(rank)

    def set(self, row_ptr, col_ind, vals):
        row_ptr, col_ind, vals = _i32(row_ptr), _i32(col_ind), _f64(vals)
        n = len(row_ptr) - 1
        self._check(load_library().phase_hip_solver_set(self._h, n, row_ptr, col_ind, vals), "set")
        self.rank = n

    def set_rhs(self, b):
        self._check(load_library().phase_hip_solver_set_rhs(self._h, _f64(b)), "setRhs")

    def set_guess(self, x0):
        self._check(load_library().phase_hip_solver_set_guess(self._h, _f64(x0)), "setGuess")

    def solve(self, raise_on_noconv=True):
        rr = ctypes.c_double()
        rc = load_library().phase_hip_solver_solve(self._h, ctypes.byref(rr))
        if rc != 0 and (raise_on_noconv or rc != -4):
            self._check(rc, "solve")
        return rr.value

    def x(self):
        out = np.empty(self.rank)
        self._check(load_library().phase_hip_solver_get_x(self._h, out), "x")
        return out

    def n_iters(self):
        return load_library().phase_hip_solver_n_iters(self._h)

    def error(self):
        return load_library().phase_hip_solver_error(self._h)

    def stats(self):
        a, b, c, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._check(load_library().phase_hip_solver_stats(self._h, a, b, c, d), "stats")
        return dict(inner_iters=a.value, outer_passes=b.value, fallbacks=c.value, solve_ms=d.value)

    def set_equation(self, eq):
        """FiniteVolumeEquation<T>::solve's hand-over: A and b = -rhs from a
        device equation, device to device."""
        self._check(load_library().phase_hip_solver_set_equation(self._h, eq._h), "setEquation")
        self.rank = eq.rank

    def x_device(self, out):
        self._check(load_library().phase_hip_solver_get_x_device(self._h, out.ptr), "x")

    def set_guess_device(self, x0):
        self._check(load_library().phase_hip_solver_set_guess_device(self._h, x0.ptr), "setGuess")

    def ordering_info(self):
        """ILU blocks and block colours of the device ordering."""
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(load_library().phase_hip_solver_ordering_info(self._h, ctypes.byref(a), ctypes.byref(b)),
                    "ordering")
        return dict(blocks=a.value, colours=b.value)

    # --- device-side helpers used by tests and the benchmark
    def export(self):
        nnz = ctypes.c_int()
        self._check(load_library().phase_hip_solver_export(self._h, nnz, None, None, None, None), "export")
        rp = np.empty(self.rank + 1, np.int32)
        ci = np.empty(nnz.value, np.int32)
        v = np.empty(nnz.value)
        rhs = np.empty(self.rank)
        self._check(load_library().phase_hip_solver_export(self._h, nnz, _p(rp), _p(ci), _p(v), _p(rhs)),
                    "export")
        return rp, ci, v, rhs

    def multigrid_levels(self):
        """Rows of each multigrid level, finest first (empty before a multigrid solve)."""
        n = ctypes.c_int(0)
        sizes = np.zeros(32, np.int32)
        self._check(load_library().phase_hip_solver_multigrid_info(self._h, ctypes.byref(n), _p(sizes), 32),
                    "multigrid_info")
        return sizes[:n.value].tolist()

    def multigrid_coarsest(self):
        """The densely inverted last multigrid level: (matrix, inverse), both
        n x n in the same row order."""
        n = self.multigrid_levels()[-1]
        a = np.zeros((n, n))
        inv = np.zeros((n, n))
        self._check(load_library().phase_hip_solver_multigrid_coarsest(self._h, n, _p(a), _p(inv)),
                    "multigrid_coarsest")
        return a, inv

    def spmv(self, x, single=False):
        y = np.empty(self.rank)
        self._check(load_library().phase_hip_solver_spmv(self._h, int(bool(single)), _f64(x), y), "spmv")
        return y

    def ilu_export(self):
        nnz = ctypes.c_int()
        self._check(load_library().phase_hip_solver_ilu_export(self._h, nnz, None, None, None, None), "ilu")
        perm = np.empty(self.rank, np.int32)
        rp = np.empty(self.rank + 1, np.int32)
        ci = np.empty(nnz.value, np.int32)
        v = np.empty(nnz.value)
        self._check(load_library().phase_hip_solver_ilu_export(self._h, nnz, _p(perm), _p(rp), _p(ci), _p(v)),
                    "ilu")
        return perm, rp, ci, v

    KERNEL_NAMES = ("spmv64", "spmv32", "lsweep64", "usweep64", "lsweep32", "usweep32", "factor64", "factor32",
                    "usolve64_launch", "lusolve64_launch", "lsolve64_launch", "spmv64_cold", "spmv32_cold",
                    "usolve32_launch", "lusolve32_launch", "lsolve32_launch", "lsolve64cg_launch",
                    "lusolve64cg_launch", "lsolve32cg_launch", "lusolve32cg_launch", "mlresid32", "mlresid64")

    def time_kernels(self, reps=20):
        """Mean device ms of each timed kernel or sweep (HIP events)."""
        out = np.zeros(len(self.KERNEL_NAMES))
        self._check(load_library().phase_hip_solver_time_kernels(self._h, int(reps), out, len(out)), "time")
        return dict(zip(self.KERNEL_NAMES, out.tolist()))

    def kernel_bytes(self):
        """Algorithmic bytes of the entries of time_kernels()."""
        out = np.zeros(len(self.KERNEL_NAMES))
        self._check(load_library().phase_hip_solver_kernel_bytes(self._h, out, len(out)), "bytes")
        return dict(zip(self.KERNEL_NAMES, out.tolist()))

    # --- learned correction network
    def learned_capture(self, every_iters, max_pairs):
        self._check(load_library().phase_hip_learned_capture(self._h, int(every_iters), int(max_pairs)), "capture")

    def learned_n_pairs(self):
        c = ctypes.c_int()
        self._check(load_library().phase_hip_learned_n_pairs(self._h, ctypes.byref(c)), "pairs")
        return c.value

    def learned_pairs(self, first=0, count=None):
        """Captured (residual, correction) pairs, original numbering: two
        float32 arrays of shape (count, rank)."""
        if count is None:
            count = self.learned_n_pairs() - first
        n = self.rank
        r = np.empty((count, n), np.float32)
        e = np.empty((count, n), np.float32)
        self._check(load_library().phase_hip_learned_pairs(self._h, int(first), int(count), _p(r), _p(e)), "pairs")
        return r, e

    def learned_add_pairs(self, r, e):
        """Append pairs in the layout learned_pairs() returns."""
        r = np.ascontiguousarray(r, dtype=np.float32)
        e = np.ascontiguousarray(e, dtype=np.float32)
        if r.shape != e.shape or r.ndim != 2 or r.shape[1] != self.rank:
            raise ValueError("pairs must be two (count, rank) arrays")
        self._check(load_library().phase_hip_learned_add_pairs(self._h, r.shape[0], _p(r), _p(e)), "addPairs")

    def learned_train(self, steps, lr=1e-3, seed=0):
        loss = ctypes.c_double()
        self._check(load_library().phase_hip_learned_train(self._h, int(steps), float(lr), int(seed),
                                                           ctypes.byref(loss)), "train")
        return loss.value

    def learned_params(self):
        c = ctypes.c_int()
        self._check(load_library().phase_hip_learned_params(self._h, None, ctypes.byref(c)), "params")
        out = np.empty(c.value, np.float32)
        self._check(load_library().phase_hip_learned_params(self._h, _p(out), ctypes.byref(c)), "params")
        return out

    def learned_set_params(self, params):
        params = np.ascontiguousarray(params, dtype=np.float32)
        self._check(load_library().phase_hip_learned_set_params(self._h, _p(params), len(params)), "setParams")

    def learned_eval(self, r, z, target):
        n = self.rank
        out = np.empty(n)
        loss = ctypes.c_double()
        grad = np.empty(len(self.learned_params()), np.float32)
        self._check(load_library().phase_hip_learned_eval(self._h, _f64(r), _f64(z), _f64(target), out,
                                                          ctypes.byref(loss), _p(grad)), "eval")
        return out, loss.value, grad

    def learned_apply(self, r):
        e = np.empty(self.rank)
        self._check(load_library().phase_hip_learned_apply(self._h, _f64(r), e), "apply")
        return e

    # --- fv:: operators on a HipMesh
    def assemble_laplacian(self, mesh, variant, gamma=1.0, gamma_face=None, gamma0_face=None, theta=1.0,
                           bc_type=None, phi_face=None, phi0_cell=None, phi0_face=None):
        args = [_f64(gamma_face), _f64(gamma0_face), _i32(bc_type), _f64(phi_face), _f64(phi0_cell),
                _f64(phi0_face)]
        rc = load_library().phase_hip_assemble_laplacian(
            self._h, mesh._h, int(variant), float(gamma), _p(args[0]), _p(args[1]), float(theta),
            _p(args[2]), _p(args[3]), _p(args[4]), _p(args[5]))
        self._check(rc, "assembleLaplacian")
        self.rank = mesh.n_cells

    def assemble_laplacian_device(self, mesh, variant, gamma=1.0, gamma_face=None, gamma0_face=None, theta=1.0,
                                  bc_type=None, phi_face=None, phi0_cell=None, phi0_face=None):
        """Same as assemble_laplacian with DeviceArray inputs (resident in HBM)."""
        rc = load_library().phase_hip_assemble_laplacian_device(
            self._h, mesh._h, int(variant), float(gamma), _dp_or_none(gamma_face), _dp_or_none(gamma0_face),
            float(theta), _dp_or_none(bc_type), _dp_or_none(phi_face), _dp_or_none(phi0_cell),
            _dp_or_none(phi0_face))
        self._check(rc, "assembleLaplacian")
        self.rank = mesh.n_cells

    def sub_rhs_device(self, v):
        self._check(load_library().phase_hip_equation_sub_rhs_device(self._h, v.ptr), "operator==")

    def zero_guess(self):
        self._check(load_library().phase_hip_solver_zero_guess(self._h), "setGuess")

    def sub_rhs(self, v):
        """CrsEquation::operator==(Vector): stored rhs -= v."""
        self._check(load_library().phase_hip_equation_sub_rhs(self._h, _f64(v)), "operator==")

    def export_padded(self):
        ns = ctypes.c_int()
        self._check(load_library().phase_hip_equation_export_padded(self._h, ns, None, None, None, None),
                    "exportPadded")
        rp = np.empty(self.rank + 1, np.int32)
        ci = np.empty(ns.value, np.int32)
        v = np.empty(ns.value)
        rhs = np.empty(self.rank)
        self._check(load_library().phase_hip_equation_export_padded(self._h, ns, _p(rp), _p(ci), _p(v),
                                                                    _p(rhs)), "exportPadded")
        return rp, ci, v, rhs
