"""ctypes front end of the CPU oracle (oracle/tng_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py, always as the checker.  The product path
(terra-ng_amd/) never imports this module.

Arrays follow the reference's unpadded Kokkos layouts:
scalar (sd, x, y, r); vector 3 x (sd, x, y, r); coords (sd, x, y, 3); radii (sd, r).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libtng_oracle.so")

NEUMANN, DIRICHLET, FREESLIP = 0, 1, 2
REPLACE, ADD = 0, 1
SKIP, COMMUNICATE = 0, 1

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = C.CDLL(_LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
        L.orc_mesh_create.restype = vp
        L.orc_mesh_create.argtypes = [ip, ip, C.c_double, C.c_double, ip, ip]
        L.orc_mesh_destroy.argtypes = [vp]
        L.orc_mesh_dims.argtypes = [vp, C.POINTER(C.c_int64)]
        for name, rt in [("orc_mesh_coords", dp), ("orc_mesh_radii", dp),
                         ("orc_mesh_boundary_mask", C.POINTER(C.c_uint8)),
                         ("orc_mesh_ownership_mask", C.POINTER(C.c_uint8)),
                         ("orc_mesh_canonical_ids", C.POINTER(C.c_int64)),
                         ("orc_mesh_subdomain_info", C.POINTER(C.c_int))]:
            getattr(L, name).restype = rt
            getattr(L, name).argtypes = [vp]
        L.orc_halo_sum.argtypes = [vp, dp, ip]
        L.orc_masked_dot.restype = C.c_double
        L.orc_masked_dot.argtypes = [vp, dp, dp, ip]
        L.orc_uniform_shell_radii.argtypes = [C.c_double, C.c_double, ip, dp]
        L.orc_trafo_nt.argtypes = [dp, dp]
        L.orc_epsdivdiv_local_matrix.argtypes = [vp, dp, ip, ip, ip, ip, ip, dp]
        L.orc_epsdivdiv_apply_slow.argtypes = [vp, dp, ip, ip, ip, ip, ip, C.c_double, dp, dp]
        L.orc_epsdivdiv_apply_fast.argtypes = [vp, dp, ip, ip, ip, ip, ip, C.c_double, dp, dp, dp]
        L.orc_penalty_modes.argtypes = [vp, dp]
        L.orc_divergence_apply.argtypes = [vp, vp, ip, ip, ip, ip, dp, dp]
        L.orc_gradient_apply.argtypes = [vp, vp, ip, ip, ip, ip, dp, dp]
        L.orc_prolongate.argtypes = [vp, vp, ip, ip, dp, dp]
        L.orc_restrict.argtypes = [vp, vp, ip, ip, dp, dp]
        L.orc_kmass_lumped_apply.argtypes = [vp, dp, dp, dp]
        L.orc_epsdivdiv_apply_stored.argtypes = [vp, dp, ip, ip, ip, ip, ip, C.c_double, dp, dp]
        L.orc_epsdivdiv_local_matrices.argtypes = [vp, dp, dp]
        L.orc_gca_elements.argtypes = [vp, dp, ip, ip, ip, C.POINTER(C.c_uint8)]
        L.orc_gca.argtypes = [vp, vp, dp, dp, ip]
        L.orc_epsdivdiv_log_k_gradient.argtypes = [vp, dp, ip, ip, dp, dp, dp]
        L.orc_vector_mass_apply.argtypes = [vp, dp, dp]
        L.orc_supg_tau.restype = C.c_double
        L.orc_supg_tau.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_supg_apply.argtypes = [vp, dp, C.c_double, C.c_double, C.c_double, ip, ip, ip, ip, ip, ip, dp, dp]
        L.orc_scalar_mass_apply.argtypes = [vp, dp, dp]
        L.orc_num_threads.restype = ip
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Mesh:
    """Oracle mesh: DistributedDomain::create_uniform(level, level_r, r_min, r_max, sdr, sdr_r)."""

    def __init__(self, lat_level, rad_level=None, r_min=0.55, r_max=1.0, lat_sdr=0, rad_sdr=0):
        L = lib()
        if rad_level is None:
            rad_level = lat_level
        self.h = L.orc_mesh_create(lat_level, rad_level, r_min, r_max, lat_sdr, rad_sdr)
        if not self.h:
            raise ValueError("invalid mesh parameters")
        d = (C.c_int64 * 5)()
        L.orc_mesh_dims(self.h, d)
        self.nsd, self.nx, self.nr, self.nalloc, self.nunique = [int(v) for v in d]
        self.shape = (self.nsd, self.nx, self.nx, self.nr)
        self.lat_level, self.rad_level = lat_level, rad_level
        self.r_min, self.r_max = r_min, r_max
        self.lat_sdr, self.rad_sdr = lat_sdr, rad_sdr

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_mesh_destroy(self.h)
            self.h = None

    def _arr(self, fn, n, dtype, shape):
        p = getattr(lib(), fn)(self.h)
        return np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True).reshape(shape)

    @property
    def coords(self):
        return self._arr("orc_mesh_coords", self.nsd * self.nx * self.nx * 3, np.float64,
                         (self.nsd, self.nx, self.nx, 3))

    @property
    def radii(self):
        return self._arr("orc_mesh_radii", self.nsd * self.nr, np.float64, (self.nsd, self.nr))

    @property
    def boundary_mask(self):
        return self._arr("orc_mesh_boundary_mask", self.nalloc, np.uint8, self.shape)

    @property
    def ownership_mask(self):
        return self._arr("orc_mesh_ownership_mask", self.nalloc, np.uint8, self.shape)

    @property
    def canonical_ids(self):
        return self._arr("orc_mesh_canonical_ids", self.nalloc, np.int64, self.shape)

    @property
    def subdomain_info(self):
        return self._arr("orc_mesh_subdomain_info", self.nsd * 4, np.int32, (self.nsd, 4))

    def halo_sum(self, data):
        a = np.ascontiguousarray(data, dtype=np.float64).copy()
        ncomp = a.size // self.nalloc
        lib().orc_halo_sum(self.h, _dp(a), ncomp)
        return a

    def masked_dot(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        return lib().orc_masked_dot(self.h, _dp(a), _dp(b), a.size // self.nalloc)

    def local_matrix(self, k, sd, xc, yc, rc, wedge):
        k = np.ascontiguousarray(k, dtype=np.float64)
        A = np.zeros((18, 18))
        lib().orc_epsdivdiv_local_matrix(self.h, _dp(k), sd, xc, yc, rc, wedge, _dp(A))
        return A

    def log_k_gradient(self, k, lam, u, bc_cmb, bc_surf):
        """d(lam^T A(k) u) / d log k per node (summed over copies)."""
        k, lam, u = (np.ascontiguousarray(a, dtype=np.float64) for a in (k, lam, u))
        out = np.zeros(self.shape)
        lib().orc_epsdivdiv_log_k_gradient(self.h, _dp(k), bc_cmb, bc_surf, _dp(lam), _dp(u), _dp(out))
        return out

    def penalty_modes(self):
        out = np.zeros((3, 3) + self.shape)
        lib().orc_penalty_modes(self.h, _dp(out))
        return out

    def epsdivdiv(self, k, src, bc_cmb, bc_surf, diagonal=False, apply_mode=REPLACE, comm=COMMUNICATE,
                  penalty_eps=1e-7, dst=None, path="slow", penalty_modes=None):
        """EpsilonDivDivKerngen::apply_impl restated (slow or fast path)."""
        k = np.ascontiguousarray(k, dtype=np.float64)
        src = np.ascontiguousarray(src, dtype=np.float64)
        out = np.zeros((3,) + self.shape) if dst is None else np.array(dst, dtype=np.float64, order="C", copy=True)
        if path == "slow":
            lib().orc_epsdivdiv_apply_slow(self.h, _dp(k), bc_cmb, bc_surf, int(diagonal), apply_mode, comm,
                                           penalty_eps, _dp(src), _dp(out))
        else:
            pm = None if penalty_modes is None else _dp(np.ascontiguousarray(penalty_modes))
            lib().orc_epsdivdiv_apply_fast(self.h, _dp(k), bc_cmb, bc_surf, int(diagonal), apply_mode, comm,
                                           penalty_eps, _dp(src), _dp(out), pm)
        return out


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def divergence(fine, coarse, u, bc_cmb, bc_surf, apply_mode=REPLACE, comm=COMMUNICATE, dst=None):
    """DivergenceKerngen::apply_impl (divergence_kerngen.hpp:206-257): fine velocity -> coarse pressure."""
    u = _c(u)
    out = np.zeros(coarse.shape) if dst is None else np.array(dst, dtype=np.float64, order="C", copy=True)
    lib().orc_divergence_apply(fine.h, coarse.h, bc_cmb, bc_surf, apply_mode, comm, _dp(u), _dp(out))
    return out


def gradient(fine, coarse, p, bc_cmb, bc_surf, apply_mode=REPLACE, comm=COMMUNICATE, dst=None):
    """GradientKerngen::apply_impl: coarse pressure -> fine velocity."""
    p = _c(p)
    out = np.zeros((3,) + fine.shape) if dst is None else np.array(dst, dtype=np.float64, order="C", copy=True)
    lib().orc_gradient_apply(fine.h, coarse.h, bc_cmb, bc_surf, apply_mode, comm, _dp(p), _dp(out))
    return out


def prolongate(coarse, fine, x, apply_mode=REPLACE, dst=None):
    """ProlongationVecConstant / scalar constant prolongation (21-point stencil)."""
    x = _c(x)
    nc = x.size // coarse.nalloc
    out = np.zeros(((nc,) if nc > 1 else ()) + fine.shape) if dst is None else np.array(dst, dtype=np.float64, order="C", copy=True)
    lib().orc_prolongate(coarse.h, fine.h, nc, apply_mode, _dp(x), _dp(out))
    return out


def restrict(fine, coarse, x, apply_mode=REPLACE, dst=None):
    """RestrictionVecConstant: owned-masked transpose stencil + additive exchange."""
    x = _c(x)
    nc = x.size // fine.nalloc
    out = np.zeros(((nc,) if nc > 1 else ()) + coarse.shape) if dst is None else np.array(dst, dtype=np.float64, order="C", copy=True)
    lib().orc_restrict(fine.h, coarse.h, nc, apply_mode, _dp(x), _dp(out))
    return out


def kmass_lumped(mesh, k, x):
    """KMass with set_lumped_diagonal(true): dst = diag(rowsum M_k) x, additive exchange."""
    k, x = _c(k), _c(x)
    out = np.zeros(mesh.shape)
    lib().orc_kmass_lumped_apply(mesh.h, _dp(k), _dp(x), _dp(out))
    return out


def local_matrices(mesh, k):
    """Neumann 18x18 local matrices of every wedge, storage layout (sd, xc, yc, rc, wedge, 18, 18)."""
    k = _c(k)
    out = np.zeros((mesh.nsd, mesh.nx - 1, mesh.nx - 1, mesh.nr - 1, 2, 18, 18))
    lib().orc_epsdivdiv_local_matrices(mesh.h, _dp(k), _dp(out))
    return out


def gca(fine, coarse, lmat_fine, exact_top=False):
    """TwoGridGCA (linalg/solvers/gca/gca.hpp:41-531): coarse stored matrices from fine ones."""
    lf = _c(lmat_fine)
    out = np.zeros((coarse.nsd, coarse.nx - 1, coarse.nx - 1, coarse.nr - 1, 2, 18, 18))
    lib().orc_gca(fine.h, coarse.h, _dp(lf), _dp(out), int(exact_top))
    return out


def gca_elements(mesh, k, level_range, coarsest):
    """GCAElementsCollector (gca_elements_collector.hpp:20-96): uint8 flags per cell of `coarsest`."""
    k = _c(k)
    out = np.zeros((coarsest.nsd, coarsest.nx - 1, coarsest.nx - 1, coarsest.nr - 1), dtype=np.uint8)
    lib().orc_gca_elements(mesh.h, _dp(k), int(level_range), coarsest.nx - 1, coarsest.nr - 1,
                           out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out


def gca_selective(fine, coarse, lmat_fine, k_coarse, level_range, flags):
    """TwoGridGCA with element selection (gca.hpp:150-175) into a Selective stored-matrix operator
    (epsilon_divdiv_kerngen.hpp:1040-1075): Galerkin matrices on wedges whose coarsest ancestor
    (cell >> level_range) is flagged, the on-the-fly Neumann matrices of k_coarse elsewhere."""
    g = gca(fine, coarse, lmat_fine)
    loc = local_matrices(coarse, k_coarse)
    nc, ncr = coarse.nx - 1, coarse.nr - 1
    xs = np.arange(nc) >> level_range
    rs = np.arange(ncr) >> level_range
    sel = np.asarray(flags, dtype=bool)[:, xs][:, :, xs][:, :, :, rs]
    return np.where(sel[..., None, None, None], g, loc)


def epsdivdiv_stored(mesh, lmat, src, bc_cmb, bc_surf, diagonal=False, apply_mode=REPLACE, comm=COMMUNICATE,
                     penalty_eps=1e-7, dst=None):
    """Slow path on stored local matrices (epsilon_divdiv_kerngen.hpp:1013-1242, stored mode Full)."""
    lm, src = _c(lmat), _c(src)
    out = np.zeros((3,) + mesh.shape) if dst is None else np.array(dst, dtype=np.float64, order="C", copy=True)
    lib().orc_epsdivdiv_apply_stored(mesh.h, _dp(lm), bc_cmb, bc_surf, int(diagonal), apply_mode, comm, penalty_eps,
                                     _dp(src), _dp(out))
    return out


def vector_mass(mesh, u):
    """VectorMass (vector_mass.hpp:66-130): consistent mass per component."""
    u = _c(u)
    out = np.zeros((3,) + mesh.shape)
    lib().orc_vector_mass_apply(mesh.h, _dp(u), _dp(out))
    return out


def supg_tau(vel_norm, kappa, h, pe_tol=1e-8):
    """supg_tau (unsteady_advection_diffusion_supg.hpp:19-49)."""
    return lib().orc_supg_tau(float(vel_norm), float(kappa), float(h), float(pe_tol))


def supg_apply(mesh, vel, src, kappa=1.0, dt=1.0, mass_scaling=1.0, treat_boundary=True, diagonal=False,
               lumped=False, supg=True, apply_mode=REPLACE, comm=COMMUNICATE, dst=None):
    """UnsteadyAdvectionDiffusionSUPG apply (unsteady_advection_diffusion_supg.hpp:214-433)."""
    vel, src = _c(vel), _c(src)
    out = np.zeros(mesh.shape) if dst is None else np.array(dst, dtype=np.float64, copy=True)
    lib().orc_supg_apply(mesh.h, _dp(vel), float(kappa), float(dt), float(mass_scaling), int(treat_boundary),
                         int(diagonal), int(lumped), int(supg), int(apply_mode), int(comm), _dp(src), _dp(out))
    return out


def scalar_mass(mesh, x):
    """Mass (mass.hpp): consistent scalar mass, additive exchange."""
    x = _c(x)
    out = np.zeros(mesh.shape)
    lib().orc_scalar_mass_apply(mesh.h, _dp(x), _dp(out))
    return out


def uniform_shell_radii(r_min, r_max, n):
    out = np.zeros(n)
    lib().orc_uniform_shell_radii(r_min, r_max, n, _dp(out))
    return out


def trafo_nt(n):
    n = np.ascontiguousarray(n, dtype=np.float64)
    R = np.zeros(9)
    lib().orc_trafo_nt(_dp(n), _dp(R))
    return R.reshape(3, 3)


def num_threads():
    return lib().orc_num_threads()
