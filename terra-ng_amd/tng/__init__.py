"""Python front end of libtng (ctypes over the C ABI in include/tng.h).

Mirrors the reference's C++ interface for the Stokes hot path so that tests
read like the reference's own: ``DistributedDomain.create_uniform``,
``VectorQ1Scalar`` / ``VectorQ1Vec``, ``EpsilonDivDivKerngen`` with
``set_operator_apply_and_communication_modes``, ``set_diagonal``,
``set_penalty_epsilon`` and ``apply``.

The library must be built (terra-ng_amd/build/libtng.so); there is no
fallback: loading fails loudly when it is missing.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "build", "libtng.so"))

# grid/shell/bit_masks.hpp BoundaryConditionFlag values
NEUMANN, DIRICHLET, FREESLIP = 0, 1, 2
# linalg/operator.hpp OperatorApplyMode / OperatorCommunicationMode
REPLACE, ADD = 0, 1
SKIP_COMMUNICATION, COMMUNICATE_ADDITIVELY = 0, 1
F64, F32 = 0, 1

# grid/shell/bit_masks.hpp ShellBoundaryFlag, grid/bit_masks.hpp NodeOwnershipFlag
INNER, CMB, SURFACE, OWNED = 1, 6, 10, 1

_STATUS = {0: "OK", 1: "INVALID_ARGUMENT", 2: "OUT_OF_MEMORY", 3: "DEVICE_ERROR", 4: "NOT_CONVERGED",
           5: "NONFINITE", 6: "UNSUPPORTED", 7: "IO_ERROR"}
INVALID_ARGUMENT, IO_ERROR = 1, 7


class TngError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{_STATUS.get(status, status)}: {msg}")
        self.status = status


class _MeshInfo(C.Structure):
    _fields_ = [("num_subdomains", C.c_int), ("nodes_lateral", C.c_int), ("nodes_radial", C.c_int),
                ("lateral_level", C.c_int), ("radial_level", C.c_int), ("lateral_sdr", C.c_int),
                ("radial_sdr", C.c_int), ("num_alloc", C.c_longlong), ("num_owned", C.c_longlong),
                ("num_interface", C.c_longlong)]


# (name, restype, argtypes) of every exported entry point of include/tng.h
_vp, _i, _d = C.c_void_p, C.c_int, C.c_double
_pv = C.POINTER(C.c_void_p)
EXPORTS = [
    ("tng_version", C.c_char_p, []),
    ("tng_last_error", C.c_char_p, [_vp]),
    ("tng_ctx_create", _i, [_i, _pv]),
    ("tng_ctx_create_with_stream", _i, [_i, _vp, _pv]),
    ("tng_ctx_destroy", _i, [_vp]),
    ("tng_ctx_synchronize", _i, [_vp]),
    ("tng_ctx_set_profiling", _i, [_vp, _i]),
    ("tng_ctx_set_tuning", _i, [_vp, C.c_char_p, _i]),
    ("tng_mesh_create_uniform", _i, [_vp, _i, _i, _d, _d, _i, _i, _pv]),
    ("tng_mesh_destroy", _i, [_vp]),
    ("tng_mesh_get_info", _i, [_vp, C.POINTER(_MeshInfo)]),
    ("tng_mesh_copy_coords", _i, [_vp, _vp]),
    ("tng_mesh_copy_radii", _i, [_vp, _vp]),
    ("tng_mesh_copy_boundary_mask", _i, [_vp, _vp]),
    ("tng_mesh_copy_ownership_mask", _i, [_vp, _vp]),
    ("tng_vec_create", _i, [_vp, _i, _i, _pv]),
    ("tng_vec_wrap", _i, [_vp, _i, _i, C.POINTER(C.c_void_p), _pv]),
    ("tng_vec_destroy", _i, [_vp]),
    ("tng_vec_upload", _i, [_vp, _vp]),
    ("tng_vec_download", _i, [_vp, _vp]),
    ("tng_vec_device_ptr", _i, [_vp, _i, _pv]),
    ("tng_vec_fill", _i, [_vp, _d]),
    ("tng_vec_copy", _i, [_vp, _vp]),
    ("tng_epsdivdiv_create", _i, [_vp, _vp, _i, _i, _i, _pv]),
    ("tng_op_set_modes", _i, [_vp, _i, _i]),
    ("tng_op_set_diagonal", _i, [_vp, _i]),
    ("tng_op_set_boundary_conditions", _i, [_vp, _i, _i]),
    ("tng_op_set_penalty_epsilon", _i, [_vp, _d]),
    ("tng_op_coefficient_updated", _i, [_vp]),
    ("tng_op_apply", _i, [_vp, _vp, _vp]),
    ("tng_op_destroy", _i, [_vp]),
    ("tng_ctx_profile_read", _i, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("tng_divergence_create", _i, [_vp, _vp, _i, _i, _pv]),
    ("tng_gradient_create", _i, [_vp, _vp, _i, _i, _pv]),
    ("tng_prolongation_create", _i, [_vp, _vp, _pv]),
    ("tng_restriction_create", _i, [_vp, _vp, _pv]),
    ("tng_kmass_lumped_create", _i, [_vp, _vp, _pv]),
    ("tng_vector_mass_create", _i, [_vp, _pv]),
    ("tng_stokes_buoyancy_rhs", _i, [_vp, _vp, _d, _vp]),
    ("tng_vec_lincomb", _i, [_vp, _i, C.POINTER(C.c_double), C.POINTER(C.c_void_p), _d]),
    ("tng_vec_scale_with_vector", _i, [_vp, _vp]),
    ("tng_vec_invert", _i, [_vp]),
    ("tng_vec_dot", _i, [_vp, _vp, C.POINTER(C.c_double)]),
    ("tng_vec_max_abs", _i, [_vp, C.POINTER(C.c_double)]),
    ("tng_vec_has_nan_or_inf", _i, [_vp, C.POINTER(C.c_int)]),
    ("tng_op_get_local_matrices", _i, [_vp, _vp]),
    ("tng_op_set_local_matrices", _i, [_vp, _vp]),
    ("tng_op_clear_local_matrices", _i, [_vp]),
    ("tng_op_copy_local_matrices", _i, [_vp, _vp]),
    ("tng_gca_assemble", _i, [_vp, _vp]),
    ("tng_gca_assemble_exact", _i, [_vp, _vp]),
    ("tng_gca_elements_collect", _i, [_vp, _i, _vp, _vp]),
    ("tng_xdmf_create", _i, [_vp, C.c_char_p, _i, _i, _pv]),
    ("tng_xdmf_add", _i, [_vp, _vp, C.c_char_p, _i]),
    ("tng_xdmf_set_write_counter", _i, [_vp, _i]),
    ("tng_xdmf_write", _i, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("tng_xdmf_destroy", _i, [_vp]),
    ("tng_xdmf_checkpoint_read", _i, [C.c_char_p, C.c_char_p, _i, _vp]),
    ("tng_gca_assemble_selective", _i, [_vp, _vp, _i, _vp, _vp]),
    ("tng_epsdivdiv_log_coefficient_gradient", _i, [_vp, _vp, _vp, _vp]),
    ("tng_stokes_adjoint_gradient", _i, [_vp, _vp, _vp, _vp, _vp]),
    ("tng_pcg_solve", _i, [_vp, _vp, _vp, _i, _d, _d, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    ("tng_power_iteration", _i, [_vp, _vp, _i, C.c_ulonglong, C.POINTER(C.c_double)]),
    ("tng_chebyshev_smooth", _i, [_vp, _vp, _i, _i, _d, _vp, _vp]),
    ("tng_jacobi_smooth", _i, [_vp, _vp, _i, _d, _vp, _vp]),
    ("tng_stokes_params_default", None, [_vp]),
    ("tng_stokes_solver_create", _i, [_vp, _i, _i, _i, _d, _d, _i, _i, _i, _i, _vp, _pv]),
    ("tng_stokes_solver_destroy", _i, [_vp]),
    ("tng_stokes_solver_num_levels", _i, [_vp, C.POINTER(C.c_int)]),
    ("tng_stokes_solver_mesh", _i, [_vp, _i, _pv]),
    ("tng_stokes_solver_set_viscosity", _i, [_vp, _vp]),
    ("tng_stokes_apply", _i, [_vp, _vp, _vp, _vp, _vp]),
    ("tng_stokes_precondition_velocity", _i, [_vp, _vp, _vp]),
    ("tng_stokes_enforce_rhs", _i, [_vp, _vp]),
    ("tng_stokes_solve", _i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("tng_vec_norm2", _i, [_vp, C.POINTER(C.c_double)]),
    ("tng_vec_dot_device", _i, [_vp, _vp, _vp]),
    ("tng_vec_fill_seeded", _i, [_vp, C.c_ulonglong]),
    ("tng_stokes_create", _i, [_vp, _vp, _vp, _i, _i, _i, _pv]),
    ("tng_stokes_op_apply", _i, [_vp, _vp, _vp, _vp, _vp]),
    ("tng_stokes_block", _i, [_vp, _i, _i, _pv]),
    ("tng_op_set_stored_matrix_mode", _i, [_vp, _i]),
    ("tng_op_get_local_matrix", _i, [_vp, _i, _i, _i, _i, _i, _vp]),
    ("tng_op_set_local_matrix", _i, [_vp, _i, _i, _i, _i, _i, _vp]),
    ("tng_epsdivdiv_apply_noalloc", _i, [_vp, _vp, _vp]),
    ("tng_supg_create", _i, [_vp, _vp, _d, _d, _i, _i, _d, _i, _pv]),
    ("tng_supg_set_dt", _i, [_vp, _d]),
    ("tng_supg_set_enabled", _i, [_vp, _i]),
    ("tng_mass_apply", _i, [_vp, _vp]),
    ("tng_energy_params_default", None, [_vp]),
    ("tng_energy_create", _i, [_vp, _vp, _vp, _vp, _pv]),
    ("tng_energy_destroy", _i, [_vp]),
    ("tng_energy_compute_dt", _i, [_vp, C.POINTER(C.c_double)]),
    ("tng_energy_step", _i, [_vp, _d, _vp]),
    ("tng_stokes_solver_update_viscosity", _i, [_vp, _vp, _d]),
    ("tng_timestep", _i, [_vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_double), _vp, _vp]),
]

class StokesParams(C.Structure):
    """apps/mantlecirculation/src/parameters.hpp:155-185 StokesSolverParameters (+ precision, coarse PCG, seed)."""
    _fields_ = [("krylov_restart", C.c_int), ("krylov_max_iterations", C.c_int),
                ("krylov_relative_tolerance", C.c_double), ("krylov_absolute_tolerance", C.c_double),
                ("viscous_pc_num_vcycles", C.c_int), ("viscous_pc_chebyshev_order", C.c_int),
                ("viscous_pc_num_smoothing_steps_prepost", C.c_int), ("viscous_pc_num_power_iterations", C.c_int),
                ("gca", C.c_int), ("viscous_pc_dtype", C.c_int), ("coarse_max_iterations", C.c_int),
                ("coarse_relative_tolerance", C.c_double), ("coarse_absolute_tolerance", C.c_double),
                ("seed", C.c_ulonglong), ("gca_coarsest_only", C.c_int), ("gca_exact", C.c_int),
                ("viscous_pc_smoother", C.c_int), ("krylov_orthogonalization", C.c_int),
                ("rotation_correction", C.c_int), ("coarse_solver", C.c_int)]

    @classmethod
    def defaults(cls, **kw):
        p = cls()
        load_library().tng_stokes_params_default(C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        return p


class EnergyParams(C.Structure):
    """apps/mantlecirculation/src/parameters.hpp EnergySolverParameters + diffusivity, temperature BCs,
    dt_scaling and energy_substeps."""
    _fields_ = [("krylov_restart", C.c_int), ("krylov_max_iterations", C.c_int),
                ("krylov_relative_tolerance", C.c_double), ("krylov_absolute_tolerance", C.c_double),
                ("diffusivity", C.c_double), ("temperature_cmb", C.c_double), ("temperature_surface", C.c_double),
                ("dt_scaling", C.c_double), ("energy_substeps", C.c_int)]

    @classmethod
    def defaults(cls, **kw):
        p = cls()
        load_library().tng_energy_params_default(C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        return p


class TimestepParams(C.Structure):
    _fields_ = [("picard_iterations", C.c_int), ("viscosity_law", C.c_int), ("rmu", C.c_double),
                ("rayleigh_number", C.c_double), ("refresh_hierarchy", C.c_int)]


class SolveReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("restarts", C.c_int), ("converged", C.c_int), ("breakdowns", C.c_int),
                ("nonfinite_preconditioner", C.c_int), ("nonfinite_residual", C.c_int),
                ("coarse_iterations", C.c_longlong), ("initial_residual", C.c_double),
                ("final_residual", C.c_double), ("relative_residual", C.c_double),
                ("history", C.POINTER(C.c_double)), ("history_capacity", C.c_int), ("history_len", C.c_int)]


_lib = None


def load_library():
    """Load libtng.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libtng.so not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, rt, args in EXPORTS:
            f = getattr(L, name)
            f.restype = rt
            f.argtypes = args
        _lib = L
    return _lib


class Context:
    """Device + stream (util/init.hpp:75-80 terra_initialize)."""

    def __init__(self, device=0):
        L = load_library()
        h = C.c_void_p()
        st = L.tng_ctx_create(device, C.byref(h))
        if st != 0:
            raise TngError(st, L.tng_last_error(None).decode())
        self.h = h

    def check(self, st):
        if st != 0:
            raise TngError(st, load_library().tng_last_error(self.h).decode())

    def synchronize(self):
        self.check(load_library().tng_ctx_synchronize(self.h))

    def set_tuning(self, name, value):
        """Launch-shape setting of the apply kernels (negative: the default)."""
        self.check(load_library().tng_ctx_set_tuning(self.h, name.encode(), int(value)))

    @contextlib.contextmanager
    def tuning(self, **settings):
        """Settings for the duration of a with-block, then the defaults."""
        try:
            for k, v in settings.items():
                self.set_tuning(k, v)
            yield self
        finally:
            for k in settings:
                self.set_tuning(k, -1)

    def set_profiling(self, on=True):
        self.check(load_library().tng_ctx_set_profiling(self.h, int(on)))

    def profile_read(self):
        """(applies, avg main-kernel ms, avg whole-apply ms) since the last read."""
        n, a, b = C.c_int(), C.c_double(), C.c_double()
        self.check(load_library().tng_ctx_profile_read(self.h, C.byref(n), C.byref(a), C.byref(b)))
        return n.value, a.value, b.value

    def close(self):
        if self.h:
            load_library().tng_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class DistributedDomain:
    """One level of the thick spherical shell (spherical_shell.hpp:2563-2711)."""

    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h
        info = _MeshInfo()
        ctx.check(load_library().tng_mesh_get_info(h, C.byref(info)))
        self.info = info
        self.nsd, self.nx, self.nr = info.num_subdomains, info.nodes_lateral, info.nodes_radial
        self.shape = (self.nsd, self.nx, self.nx, self.nr)
        self.num_alloc, self.num_owned = info.num_alloc, info.num_owned

    @classmethod
    def create_uniform(cls, lateral_level, radial_level, r_min, r_max, lateral_sdr=0, radial_sdr=0, ctx=None):
        ctx = ctx or default_context()
        h = C.c_void_p()
        ctx.check(load_library().tng_mesh_create_uniform(ctx.h, lateral_level, radial_level, r_min, r_max,
                                                         lateral_sdr, radial_sdr, C.byref(h)))
        return cls(ctx, h)

    def _copy(self, fn, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        self.ctx.check(getattr(load_library(), fn)(self.h, out.ctypes.data_as(C.c_void_p)))
        return out

    def coords(self):
        return self._copy("tng_mesh_copy_coords", (self.nsd, self.nx, self.nx, 3), np.float64)

    def radii(self):
        return self._copy("tng_mesh_copy_radii", (self.nsd, self.nr), np.float64)

    def boundary_mask(self):
        return self._copy("tng_mesh_copy_boundary_mask", self.shape, np.uint8)

    def ownership_mask(self):
        return self._copy("tng_mesh_copy_ownership_mask", self.shape, np.uint8)

    def __del__(self):
        try:
            if self.h:
                load_library().tng_mesh_destroy(self.h)
                self.h = None
        except Exception:
            pass


class _Vector:
    ncomp = 1

    def __init__(self, domain, dtype=F64, component_ptrs=None, keepalive=None):
        self.domain, self.dtype = domain, dtype
        h = C.c_void_p()
        if component_ptrs is None:
            domain.ctx.check(load_library().tng_vec_create(domain.h, dtype, self.ncomp, C.byref(h)))
        else:
            arr = (C.c_void_p * self.ncomp)(*[int(p) for p in component_ptrs])
            domain.ctx.check(load_library().tng_vec_wrap(domain.h, dtype, self.ncomp, arr, C.byref(h)))
        self._keepalive = keepalive
        self.h = h

    @classmethod
    def wrap(cls, domain, dtype, component_ptrs, keepalive=None):
        """Non-owning view of existing device arrays (one per component)."""
        return cls(domain, dtype, component_ptrs, keepalive)

    @property
    def np_dtype(self):
        return np.float64 if self.dtype == F64 else np.float32

    @property
    def host_shape(self):
        return ((self.ncomp,) if self.ncomp > 1 else ()) + self.domain.shape

    def upload(self, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=self.np_dtype).reshape(self.host_shape))
        self.domain.ctx.check(load_library().tng_vec_upload(self.h, a.ctypes.data_as(C.c_void_p)))
        return self

    def download(self):
        out = np.empty(self.host_shape, dtype=self.np_dtype)
        self.domain.ctx.check(load_library().tng_vec_download(self.h, out.ctypes.data_as(C.c_void_p)))
        return out

    def fill(self, value):
        self.domain.ctx.check(load_library().tng_vec_fill(self.h, float(value)))

    def assign(self, other):
        self.domain.ctx.check(load_library().tng_vec_copy(self.h, other.h))

    def device_ptr(self, comp=0):
        p = C.c_void_p()
        self.domain.ctx.check(load_library().tng_vec_device_ptr(self.h, comp, C.byref(p)))
        return p.value

    def __del__(self):
        try:
            if self.h:
                load_library().tng_vec_destroy(self.h)
                self.h = None
        except Exception:
            pass


class VectorQ1Scalar(_Vector):
    """linalg/vector_q1.hpp:19-155"""
    ncomp = 1


class VectorQ1Vec(_Vector):
    """linalg/vector_q1.hpp:166-304"""
    ncomp = 3


class _Operator:
    """Common part of the operators: modes, apply, lifetime (linalg/operator.hpp:29-159)."""

    def __init__(self, domain, h):
        self.domain, self.h = domain, h

    def set_operator_apply_and_communication_modes(self, apply_mode, comm_mode):
        self.domain.ctx.check(load_library().tng_op_set_modes(self.h, apply_mode, comm_mode))

    def apply(self, src, dst):
        self.domain.ctx.check(load_library().tng_op_apply(self.h, src.h, dst.h))

    def __del__(self):
        try:
            if self.h:
                load_library().tng_op_destroy(self.h)
                self.h = None
        except Exception:
            pass


def _create(domain, fn, *args):
    h = C.c_void_p()
    domain.ctx.check(getattr(load_library(), fn)(*args, C.byref(h)))
    return h


class EpsilonDivDivKerngen(_Operator):
    """fe/wedge/operators/shell/epsilon_divdiv_kerngen.hpp:70-2393 (the HIP path)."""

    def __init__(self, domain, k, bcs=(DIRICHLET, DIRICHLET), diagonal=False):
        super().__init__(domain, _create(domain, "tng_epsdivdiv_create", domain.h, k.h, int(bcs[0]), int(bcs[1]),
                                         int(diagonal)))
        self.k = k

    def set_diagonal(self, v):
        self.domain.ctx.check(load_library().tng_op_set_diagonal(self.h, int(v)))

    def set_boundary_conditions(self, bcs):
        self.domain.ctx.check(load_library().tng_op_set_boundary_conditions(self.h, int(bcs[0]), int(bcs[1])))

    def set_penalty_epsilon(self, eps):
        self.domain.ctx.check(load_library().tng_op_set_penalty_epsilon(self.h, float(eps)))

    def coefficient_updated(self):
        self.domain.ctx.check(load_library().tng_op_coefficient_updated(self.h))

    def log_coefficient_gradient(self, lam, u, g):
        """g = d(lam^T A(k) u)/d log k per node (no reference counterpart)."""
        self.domain.ctx.check(load_library().tng_epsdivdiv_log_coefficient_gradient(self.h, lam.h, u.h, g.h))

    def _lmat_shape(self):
        d = self.domain
        return (d.nsd, d.nx - 1, d.nx - 1, d.nr - 1, 2, 18, 18)

    def get_local_matrices(self):
        """Stored matrices, or the Neumann local matrices assembled from k."""
        out = np.empty(self._lmat_shape())
        self.domain.ctx.check(load_library().tng_op_get_local_matrices(self.h, out.ctypes.data_as(C.c_void_p)))
        return out

    def set_local_matrices(self, lmat):
        a = np.ascontiguousarray(lmat, dtype=np.float64).reshape(self._lmat_shape())
        self.domain.ctx.check(load_library().tng_op_set_local_matrices(self.h, a.ctypes.data_as(C.c_void_p)))

    def clear_local_matrices(self):
        self.domain.ctx.check(load_library().tng_op_clear_local_matrices(self.h))

    def copy_local_matrices(self, other):
        self.domain.ctx.check(load_library().tng_op_copy_local_matrices(self.h, other.h))

    def set_stored_matrix_mode(self, full):
        """OperatorStoredMatrixMode::Full (zero matrices) or Off."""
        self.domain.ctx.check(load_library().tng_op_set_stored_matrix_mode(self.h, int(bool(full))))

    def get_local_matrix(self, sd, x_cell, y_cell, r_cell, wedge):
        """GCACapable get_local_matrix: an 18 x 18 array."""
        out = np.empty((18, 18))
        self.domain.ctx.check(load_library().tng_op_get_local_matrix(self.h, sd, x_cell, y_cell, r_cell, wedge,
                                                                     out.ctypes.data_as(C.c_void_p)))
        return out

    def set_local_matrix(self, sd, x_cell, y_cell, r_cell, wedge, mat):
        a = np.ascontiguousarray(mat, dtype=np.float64).reshape(18, 18)
        self.domain.ctx.check(load_library().tng_op_set_local_matrix(self.h, sd, x_cell, y_cell, r_cell, wedge,
                                                                     a.ctypes.data_as(C.c_void_p)))


class _Block(_Operator):
    """A block of a Stokes operator: a borrowed handle that the Stokes operator releases."""

    def __init__(self, domain, h, owner):
        super().__init__(domain, h)
        self.owner = owner

    def __del__(self):
        self.h = None


class EpsDivDivStokes:
    """fe/wedge/operators/shell/epsilon_divdiv_stokes.hpp:18-133: K = [A B^T; B 0]."""

    def __init__(self, domain_fine, domain_coarse, k, bcs=(DIRICHLET, DIRICHLET), diagonal=False):
        self.domain, self.domain_coarse, self.k = domain_fine, domain_coarse, k
        self.h = _create(domain_fine, "tng_stokes_create", domain_fine.h, domain_coarse.h, k.h, int(bcs[0]),
                         int(bcs[1]), int(diagonal))

    def apply(self, u, p, out_u, out_p):
        self.domain.ctx.check(load_library().tng_stokes_op_apply(self.h, u.h, p.h, out_u.h, out_p.h))

    def _block(self, row, col):
        h = C.c_void_p()
        self.domain.ctx.check(load_library().tng_stokes_block(self.h, row, col, C.byref(h)))
        return _Block(self.domain if col == 1 or row == 1 else self.domain_coarse, h, self)

    def block_11(self):
        return self._block(1, 1)

    def block_12(self):
        return self._block(1, 2)

    def block_21(self):
        return self._block(2, 1)

    def block_22(self):
        return self._block(2, 2)

    def __del__(self):
        try:
            if self.h:
                load_library().tng_op_destroy(self.h)
                self.h = None
        except Exception:
            pass


def gca_assemble(fine_op, coarse_op, exact=False):
    """TwoGridGCA (linalg/solvers/gca/gca.hpp): coarse stored matrices from the fine operator.
    exact=True forms the Galerkin product R A P instead of the reference's top-face rule."""
    lib = load_library()
    if exact:
        coarse_op.domain.ctx.check(lib.tng_gca_assemble_exact(fine_op.h, coarse_op.h))
    else:
        coarse_op.domain.ctx.check(lib.tng_gca_assemble(fine_op.h, coarse_op.h))


def _cell_shape(domain):
    return (domain.nsd, domain.nx - 1, domain.nx - 1, domain.nr - 1)


def gca_elements_collect(k, level_range, domain_coarsest):
    """GCAElementsCollector (linalg/solvers/gca/gca_elements_collector.hpp:20-96): uint8 flags per
    coarsest-level cell (sd, x_cell, y_cell, r_cell), 1 where a descendant wedge on k's mesh has a
    reference-coordinate |grad k| above 10 at its centroid."""
    out = np.zeros(_cell_shape(domain_coarsest), dtype=np.uint8)
    domain_coarsest.ctx.check(load_library().tng_gca_elements_collect(
        k.h, int(level_range), domain_coarsest.h, out.ctypes.data_as(C.c_void_p)))
    return out


def gca_assemble_selective(fine_op, coarse_op, level_range, domain_coarsest, flags):
    """TwoGridGCA with element selection (gca.hpp:150-175) into a Selective stored-matrix operator:
    Galerkin matrices where the coarsest ancestor is flagged, re-discretised ones elsewhere."""
    f = np.ascontiguousarray(flags, dtype=np.uint8)
    if f.shape != _cell_shape(domain_coarsest):
        raise ValueError(f"flags must have shape {_cell_shape(domain_coarsest)}, got {f.shape}")
    coarse_op.domain.ctx.check(load_library().tng_gca_assemble_selective(
        fine_op.h, coarse_op.h, int(level_range), domain_coarsest.h, f.ctypes.data_as(C.c_void_p)))


class XDMFOutput:
    """io/xdmf.hpp:148-513 XDMFOutput: node fields of one mesh written for XDMF readers and as checkpoints.

    Fields are added once (``add``), then every ``write()`` writes one step; the first also writes the
    geometry, the wedge topology and the checkpoint metadata."""

    def __init__(self, domain, directory, points_bytes=4, connectivity_bytes=8):
        self.domain = domain
        self._fields = []  # keeps the added vectors alive
        h = C.c_void_p()
        domain.ctx.check(load_library().tng_xdmf_create(domain.h, os.fsencode(directory), int(points_bytes),
                                                        int(connectivity_bytes), C.byref(h)))
        self.h = h

    def add(self, field, label, output_bytes=4):
        self.domain.ctx.check(load_library().tng_xdmf_add(self.h, field.h, label.encode(), int(output_bytes)))
        self._fields.append(field)

    def set_write_counter(self, counter):
        self.domain.ctx.check(load_library().tng_xdmf_set_write_counter(self.h, int(counter)))

    def write(self):
        """Writes one step; returns (device seconds of the layout kernels, wall seconds of the call)."""
        pack, total = C.c_double(), C.c_double()
        self.domain.ctx.check(load_library().tng_xdmf_write(self.h, C.byref(pack), C.byref(total)))
        return pack.value, total.value

    def close(self):
        if getattr(self, "h", None):
            load_library().tng_xdmf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_xdmf_checkpoint_metadata(directory):
    """io/xdmf.hpp:1060-1186 read_xdmf_checkpoint_metadata: the checkpoint's structure as a dict
    (version, subdomains per diamond side / radially, subdomain sizes, radii, grid scalar bytes,
    grid data files, subdomain ordering and the global subdomain ids in file order).  Host only."""
    import struct
    path = os.path.join(directory, "checkpoint_metadata.bin")
    if not os.path.exists(path):
        raise TngError(IO_ERROR, f"Could not find file: {path}")
    with open(path, "rb") as f:
        b = f.read()
    pos = 0

    def take(fmt):
        nonlocal pos
        size = struct.calcsize("<" + fmt)
        if pos + size > len(b):
            raise TngError(IO_ERROR, "Failed to read from input stream.")
        v = struct.unpack_from("<" + fmt, b, pos)
        pos += size
        return v if len(v) > 1 else v[0]

    md = {"version": take("i")}
    md["num_subdomains_per_diamond_lateral_direction"] = take("i")
    md["num_subdomains_per_diamond_radial_direction"] = take("i")
    md["size_x"], md["size_y"], md["size_r"] = take("i"), take("i"), take("i")
    nrad = md["num_subdomains_per_diamond_radial_direction"] * (md["size_r"] - 1) + 1
    radii = [take("d") for _ in range(nrad)]
    if any(radii[i] <= radii[i - 1] for i in range(1, nrad)):
        raise TngError(IO_ERROR, "Radii are not sorted correctly in checkpoint.")
    md["radii"] = radii
    md["grid_scalar_bytes"] = take("i") if md["version"] > 0 else 0
    files = []
    for _ in range(take("i")):
        n = take("i")
        if pos + n > len(b):
            raise TngError(IO_ERROR, "Failed to read from input stream.")
        name = b[pos:pos + n].decode()
        pos += n
        dt, sb, vd = take("i"), take("i"), take("i")
        files.append({"grid_name_string": name, "scalar_data_type": dt, "scalar_bytes": sb, "vec_dim": vd})
    md["grid_data_files"] = files
    md["checkpoint_subdomain_ordering"] = take("i")
    if md["checkpoint_subdomain_ordering"] == 0:
        s_lat = md["num_subdomains_per_diamond_lateral_direction"]
        n_ids = 10 * s_lat * s_lat * md["num_subdomains_per_diamond_radial_direction"]
        md["checkpoint_ordering_0_global_subdomain_ids"] = [take("q") for _ in range(n_ids)]
    return md


def read_xdmf_checkpoint(directory, label, step, out):
    """io/xdmf.hpp:1199-1441 read_xdmf_checkpoint_grid into the vector ``out`` (its mesh must match)."""
    out.domain.ctx.check(load_library().tng_xdmf_checkpoint_read(os.fsencode(directory), label.encode(), int(step),
                                                                 out.h))


class DivergenceKerngen(_Operator):
    """fe/wedge/operators/shell/divergence_kerngen.hpp:121-257: fine velocity -> coarse pressure."""

    def __init__(self, domain_fine, domain_coarse, bcs=(DIRICHLET, DIRICHLET)):
        super().__init__(domain_fine, _create(domain_fine, "tng_divergence_create", domain_fine.h, domain_coarse.h,
                                              int(bcs[0]), int(bcs[1])))
        self.domain_coarse = domain_coarse


class GradientKerngen(_Operator):
    """fe/wedge/operators/shell/gradient_kerngen.hpp:126-258: coarse pressure -> fine velocity."""

    def __init__(self, domain_fine, domain_coarse, bcs=(DIRICHLET, DIRICHLET)):
        super().__init__(domain_fine, _create(domain_fine, "tng_gradient_create", domain_fine.h, domain_coarse.h,
                                              int(bcs[0]), int(bcs[1])))
        self.domain_coarse = domain_coarse


class ProlongationConstant(_Operator):
    """fe/wedge/operators/shell/prolongation_constant.hpp:14-200 (scalar or vector by the vectors)."""

    def __init__(self, domain_coarse, domain_fine):
        super().__init__(domain_fine, _create(domain_fine, "tng_prolongation_create", domain_coarse.h,
                                              domain_fine.h))
        self.domain_coarse = domain_coarse


ProlongationVecConstant = ProlongationConstant


class RestrictionConstant(_Operator):
    """fe/wedge/operators/shell/restriction_constant.hpp:18-230 (scalar or vector by the vectors)."""

    def __init__(self, domain_fine, domain_coarse):
        super().__init__(domain_fine, _create(domain_fine, "tng_restriction_create", domain_fine.h,
                                              domain_coarse.h))
        self.domain_coarse = domain_coarse


RestrictionVecConstant = RestrictionConstant


class KMassLumped(_Operator):
    """fe/wedge/operators/shell/kmass.hpp:44-190 with lumped_diagonal = true."""

    def __init__(self, domain, k):
        super().__init__(domain, _create(domain, "tng_kmass_lumped_create", domain.h, k.h))
        self.k = k


class UnsteadyAdvectionDiffusionSUPG(_Operator):
    """fe/wedge/operators/shell/unsteady_advection_diffusion_supg.hpp:128-433 (scalar, FP64)."""

    def __init__(self, domain, velocity, diffusivity, dt, treat_boundary=True, diagonal=False, mass_scaling=1.0,
                 lumped_mass=False):
        super().__init__(domain, _create(domain, "tng_supg_create", domain.h, velocity.h, float(diffusivity),
                                         float(dt), int(treat_boundary), int(diagonal), float(mass_scaling),
                                         int(lumped_mass)))
        self.velocity = velocity

    def set_dt(self, dt):
        self.domain.ctx.check(load_library().tng_supg_set_dt(self.h, float(dt)))

    def set_supg_enabled(self, on):
        self.domain.ctx.check(load_library().tng_supg_set_enabled(self.h, int(bool(on))))


def mass_apply(src, dst):
    """Mass (mass.hpp): dst = M src, consistent scalar mass."""
    src.domain.ctx.check(load_library().tng_mass_apply(src.h, dst.h))


class VectorMass(_Operator):
    """fe/wedge/operators/shell/vector_mass.hpp:17-130 (consistent, per component)."""

    def __init__(self, domain):
        super().__init__(domain, _create(domain, "tng_vector_mass_create", domain.h))


# ------------------------------------------------- linalg (linalg/vector.hpp)

def lincomb(y, c, x, c0=0.0):
    """y = sum c[i] x[i] + c0 (linalg/vector.hpp:72-95)."""
    n = len(c)
    if n != len(x):
        raise ValueError("lincomb: c and x must have the same size")
    ca = (C.c_double * max(n, 1))(*[float(v) for v in c])
    xa = (C.c_void_p * max(n, 1))(*[v.h.value if isinstance(v.h, C.c_void_p) else v.h for v in x])
    y.domain.ctx.check(load_library().tng_vec_lincomb(y.h, n, ca, xa, float(c0)))


def scale_in_place(y, x):
    y.domain.ctx.check(load_library().tng_vec_scale_with_vector(y.h, x.h))


def invert_entries(y):
    y.domain.ctx.check(load_library().tng_vec_invert(y.h))


def dot(x, y):
    out = C.c_double()
    x.domain.ctx.check(load_library().tng_vec_dot(x.h, y.h, C.byref(out)))
    return out.value


def norm_2(x):
    out = C.c_double()
    x.domain.ctx.check(load_library().tng_vec_norm2(x.h, C.byref(out)))
    return out.value


def dot_device(x, y, out_device_ptr):
    """dot(x, y) written to a device address (8 bytes), without a host synchronisation."""
    x.domain.ctx.check(load_library().tng_vec_dot_device(x.h, y.h, C.c_void_p(out_device_ptr)))


def randomize(x, seed=12345):
    """Uniform [0, 1) entries, reproducible from the seed (linalg/vector.hpp randomize_impl)."""
    x.domain.ctx.check(load_library().tng_vec_fill_seeded(x.h, C.c_ulonglong(seed)))


def norm_inf(x):
    out = C.c_double()
    x.domain.ctx.check(load_library().tng_vec_max_abs(x.h, C.byref(out)))
    return out.value


def pcg_solve(op, x, b, max_iterations=50, rtol=1e-6, atol=1e-16):
    """PCG with the identity preconditioner (linalg/solvers/pcg.hpp); returns (iterations, relative residual)."""
    it, rel = C.c_int(), C.c_double()
    op.domain.ctx.check(load_library().tng_pcg_solve(op.h, x.h, b.h, max_iterations, rtol, atol, C.byref(it),
                                                     C.byref(rel)))
    return it.value, rel.value


def power_iteration(op, inv_diag, iterations=10, seed=12345):
    """Largest-eigenvalue estimate of D^{-1} A (linalg/solvers/power_iteration.hpp)."""
    out = C.c_double()
    op.domain.ctx.check(load_library().tng_power_iteration(op.h, inv_diag.h, iterations, seed, C.byref(out)))
    return out.value


def chebyshev_smooth(op, inv_diag, x, b, order=2, iterations=1, lambda_max_estimate=1.0):
    """Chebyshev sweeps with a given eigenvalue estimate (linalg/solvers/chebyshev.hpp)."""
    op.domain.ctx.check(load_library().tng_chebyshev_smooth(op.h, inv_diag.h, order, iterations,
                                                            float(lambda_max_estimate), x.h, b.h))


def jacobi_smooth(op, inv_diag, x, b, iterations=1, omega=1.0):
    """Damped Jacobi sweeps x += omega D^{-1} (b - A x) (linalg/solvers/jacobi.hpp)."""
    op.domain.ctx.check(load_library().tng_jacobi_smooth(op.h, inv_diag.h, iterations, float(omega), x.h, b.h))


class _BorrowedDomain(DistributedDomain):
    """A mesh owned by a solver (not destroyed from Python).  It keeps the
    solver alive while vectors on it exist; the solver does not hold it, so
    no reference cycle delays or reorders their destruction."""

    def __init__(self, ctx, h, owner):
        super().__init__(ctx, h)
        self._owner = owner

    def __del__(self):
        pass


class StokesSolver:
    """StokesContext (apps/mantlecirculation/src/stokes_solver.hpp:143-691) for one GPU."""

    def __init__(self, min_level, max_level, r_min=0.55, r_max=1.0, bcs=(DIRICHLET, DIRICHLET), params=None,
                 radial_extra_levels=0, lateral_sdr=0, radial_sdr=0, ctx=None):
        self.ctx = ctx or default_context()
        self.params = params if params is not None else StokesParams.defaults()
        h = C.c_void_p()
        self.ctx.check(load_library().tng_stokes_solver_create(
            self.ctx.h, min_level, max_level, radial_extra_levels, r_min, r_max, lateral_sdr, radial_sdr,
            int(bcs[0]), int(bcs[1]), C.byref(self.params), C.byref(h)))
        self.h = h
        n = C.c_int()
        self.ctx.check(load_library().tng_stokes_solver_num_levels(h, C.byref(n)))
        self.num_levels = n.value

    def domain(self, level):
        """Mesh of multigrid level `level` (0 = coarsest)."""
        mh = C.c_void_p()
        self.ctx.check(load_library().tng_stokes_solver_mesh(self.h, level, C.byref(mh)))
        return _BorrowedDomain(self.ctx, mh, self)

    @property
    def domain_velocity(self):
        return self.domain(self.num_levels - 1)

    @property
    def domain_pressure(self):
        return self.domain(self.num_levels - 2)

    def set_viscosity(self, eta):
        self.ctx.check(load_library().tng_stokes_solver_set_viscosity(self.h, eta.h))

    def apply(self, u, p, out_u, out_p):
        self.ctx.check(load_library().tng_stokes_apply(self.h, u.h, p.h, out_u.h, out_p.h))

    def precondition_velocity(self, x, b):
        self.ctx.check(load_library().tng_stokes_precondition_velocity(self.h, x.h, b.h))

    def buoyancy_rhs(self, temperature, rayleigh, f_u):
        """f_u = M (Ra T x/|x|), boundary-enforced (StokesContext::solve's RHS)."""
        self.ctx.check(load_library().tng_stokes_buoyancy_rhs(self.h, temperature.h, float(rayleigh), f_u.h))

    def enforce_rhs(self, f_u):
        self.ctx.check(load_library().tng_stokes_enforce_rhs(self.h, f_u.h))

    def solve(self, u, p, f_u, f_p, history_capacity=1000):
        rep = SolveReport()
        hist = (C.c_double * history_capacity)()
        rep.history = C.cast(hist, C.POINTER(C.c_double))
        rep.history_capacity = history_capacity
        self.ctx.check(load_library().tng_stokes_solve(self.h, u.h, p.h, f_u.h, f_p.h, C.byref(rep)))
        out = {name: getattr(rep, name) for name, _ in SolveReport._fields_ if name not in ("history",)}
        out["history"] = [hist[i] for i in range(min(rep.history_len, history_capacity))]
        return out

    def adjoint_gradient(self, u, dJdu, g, history_capacity=1000):
        """g = dJ/dlog(eta) on the velocity mesh for J(u), u the solution of the last solve."""
        rep = SolveReport()
        hist = (C.c_double * history_capacity)()
        rep.history = C.cast(hist, C.POINTER(C.c_double))
        rep.history_capacity = history_capacity
        self.ctx.check(load_library().tng_stokes_adjoint_gradient(self.h, u.h, dJdu.h, g.h, C.byref(rep)))
        return {name: getattr(rep, name) for name, _ in SolveReport._fields_ if name != "history"}

    def close(self):
        if self.h:
            load_library().tng_stokes_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def has_nan_or_inf(x):
    out = C.c_int()
    x.domain.ctx.check(load_library().tng_vec_has_nan_or_inf(x.h, C.byref(out)))
    return bool(out.value)


def apply(op, src, dst):
    """linalg::apply (linalg/operator.hpp:155-159)."""
    op.apply(src, dst)


class SUPGEnergySolver:
    """SUPGSolver (apps/mantlecirculation/src/energy_solver.hpp:85-276) on a Stokes solver's velocity mesh."""

    def __init__(self, stokes, velocity, temperature, params=None):
        self.stokes, self.velocity, self.temperature = stokes, velocity, temperature
        self.params = params if params is not None else EnergyParams.defaults()
        h = C.c_void_p()
        stokes.ctx.check(load_library().tng_energy_create(stokes.h, velocity.h, temperature.h, C.byref(self.params),
                                                          C.byref(h)))
        self.h = h

    def compute_dt(self):
        out = C.c_double()
        self.stokes.ctx.check(load_library().tng_energy_compute_dt(self.h, C.byref(out)))
        return out.value

    def step(self, dt, history_capacity=1000):
        rep = SolveReport()
        hist = (C.c_double * history_capacity)()
        rep.history = C.cast(hist, C.POINTER(C.c_double))
        rep.history_capacity = history_capacity
        self.stokes.ctx.check(load_library().tng_energy_step(self.h, float(dt), C.byref(rep)))
        return {name: getattr(rep, name) for name, _ in SolveReport._fields_ if name != "history"}

    def __del__(self):
        try:
            if self.h:
                load_library().tng_energy_destroy(self.h)
                self.h = None
        except Exception:
            pass


def update_viscosity(stokes, temperature, rmu):
    """StokesContext::update_viscosity with the Frank-Kamenetskii law (velocity level only)."""
    stokes.ctx.check(load_library().tng_stokes_solver_update_viscosity(stokes.h, temperature.h, float(rmu)))


def timestep(stokes, energy, u, p, picard_iterations=1, viscosity_law=0, rmu=1.0, rayleigh_number=1e5,
             refresh_hierarchy=False):
    """One step of the mantle-circulation time loop (mantlecirculation.cpp:413-445); returns
    (dt, Stokes report, energy report).  refresh_hierarchy: after the viscosity update,
    also rebuild the coarse levels (the reference updates the velocity level only)."""
    prm = TimestepParams(picard_iterations, viscosity_law, float(rmu), float(rayleigh_number),
                         int(bool(refresh_hierarchy)))
    dt = C.c_double()
    reps = []
    for _ in range(2):
        rep = SolveReport()
        rep.history = None
        rep.history_capacity = 0
        reps.append(rep)
    stokes.ctx.check(load_library().tng_timestep(stokes.h, energy.h, u.h, p.h, C.byref(prm), C.byref(dt),
                                                 C.byref(reps[0]), C.byref(reps[1])))
    conv = [{name: getattr(r, name) for name, _ in SolveReport._fields_ if name != "history"} for r in reps]
    return dt.value, conv[0], conv[1]
