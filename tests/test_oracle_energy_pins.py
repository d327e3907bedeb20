"""Pins of the oracle's energy-equation operator (oracle/tng_oracle.c
orc_supg_apply, orc_supg_tau, orc_scalar_mass_apply; reference
fe/wedge/operators/shell/unsteady_advection_diffusion_supg.hpp, mass.hpp)
against what the mathematics fixes, independent of the restated formulas:

- supg_tau: its limits (tau -> h / (2|u|) for large Peclet numbers, the
  expansion h Pe / (6|u|) for small ones, zero at zero velocity or below the
  Peclet tolerance, the 10 h / |u| cap);
- the isoparametric wedge map reproduces linear fields and the 6-point rule
  integrates every term below exactly, so
    * mass: M 1 sums (over unique nodes) to the mesh volume, computed here
      from the flat corner triangles as sum |det[p1-p0, p2-p0, p0]| / 2 *
      (r_2^3 - r_1^3) / 3;
    * diffusion: K (a . x) = 0 at every interior node (divergence theorem);
    * advection by a constant velocity c: C (a . x) = (c . a) M 1 at every
      node, and C 1 = 0;
    * streamline diffusion with constant c: zero at interior nodes for a
      linear field;
- treated boundary rows are the diagonal times the value, and the diagonal
  mode equals the diagonal of unit applies (brute force on a small mesh)."""
import numpy as np
import pytest

import oracle as O
from conftest import consistent


def _x(om):
    """Node positions (sd, x, y, r, 3)."""
    return om.coords[:, :, :, None, :] * om.radii[:, None, None, :, None]


def _interior(om):
    return (om.boundary_mask == 1) & (om.ownership_mask == 1)


def test_supg_tau_limits():
    h, kappa = 0.1, 1e-3
    for v in (10.0, 100.0, 1e4):
        pe = v * h / (2 * kappa)
        assert pe > 400
        # large Peclet: coth(Pe) - 1/Pe -> 1 - 1/Pe
        assert abs(O.supg_tau(v, kappa, h) - (h / (2 * v)) * (1 - 1 / pe)) < 1e-12 * h / v
    for v in (1e-2, 1e-1):
        pe = v * h / (2 * 1.0)
        # small Peclet: coth x - 1/x = x/3 - x^3/45 + 2 x^5/945 - ...; the
        # difference 1/tanh(x) - 1/x cancels to a relative error of about
        # 3 eps / x^2, which bounds the agreement
        ref = (h / (2 * v)) * (pe / 3 - pe ** 3 / 45 + 2 * pe ** 5 / 945)
        tol = 30 * np.finfo(float).eps / pe ** 2 + 1e-12
        assert abs(O.supg_tau(v, 1.0, h) - ref) < tol * ref
    assert O.supg_tau(0.0, kappa, h) == 0.0
    assert O.supg_tau(1e-13, kappa, h) == 0.0
    assert O.supg_tau(1e-9, 1.0, h, pe_tol=1e-8) == 0.0  # Pe = 5e-11 below the tolerance
    # never above the cap 10 h / |u|
    for v in (1e-3, 1.0, 1e3):
        assert O.supg_tau(v, 1e-9, h) <= 10 * h / v + 1e-15


@pytest.mark.parametrize("L,ls", [(2, 0), (2, 1)])
def test_mass_sums_to_mesh_volume(L, ls):
    om = O.Mesh(L, L, 0.55, 1.0, ls, 0)
    zero = np.zeros((3,) + om.shape)
    m1 = O.supg_apply(om, zero, np.ones(om.shape), kappa=0.0, dt=0.0, treat_boundary=False)
    assert np.allclose(m1, O.scalar_mass(om, np.ones(om.shape)), rtol=1e-13, atol=0)
    total = m1[om.ownership_mask == 1].sum()
    vol = 0.0
    for sd in range(om.nsd):
        p = om.coords[sd]
        r = om.radii[sd]
        shell = ((r[1:] ** 3 - r[:-1] ** 3) / 3.0).sum()
        for x in range(om.nx - 1):
            for y in range(om.nx - 1):
                for tri in (((0, 0), (1, 0), (0, 1)), ((1, 1), (0, 1), (1, 0))):
                    p0, p1, p2 = (p[x + a, y + b] for a, b in tri)
                    vol += abs(np.linalg.det(np.stack([p1 - p0, p2 - p0, p0], axis=1))) / 2.0 * shell
    assert abs(total - vol) < 1e-12 * vol


def test_diffusion_kills_linear_fields_inside():
    om = O.Mesh(2, 2)
    X = _x(om)
    zero = np.zeros((3,) + om.shape)
    inner = _interior(om)
    for a in ([1.0, 0, 0], [0.3, -0.7, 0.2]):
        T = X @ np.array(a)
        K = O.supg_apply(om, zero, T, kappa=1.0, dt=1.0, mass_scaling=0.0, treat_boundary=False)
        scale = np.abs(O.supg_apply(om, zero, np.random.default_rng(1).uniform(-1, 1, om.shape), kappa=1.0, dt=1.0,
                                    mass_scaling=0.0, treat_boundary=False)).max()
        assert np.abs(K[inner]).max() < 1e-12 * scale
        # constants too (every node)
        K1 = O.supg_apply(om, zero, np.ones(om.shape), kappa=1.0, dt=1.0, mass_scaling=0.0, treat_boundary=False)
        assert np.abs(K1).max() < 1e-12 * scale


@pytest.mark.parametrize("supg", [False, True])
def test_constant_advection_of_linear_fields(supg):
    om = O.Mesh(2, 2)
    X = _x(om)
    c = np.array([0.4, -1.1, 0.7])
    vel = np.broadcast_to(c[:, None, None, None, None], (3,) + om.shape).copy()
    a = np.array([0.5, 0.25, -1.0])
    T = X @ a
    C = O.supg_apply(om, vel, T, kappa=0.0, dt=1.0, mass_scaling=0.0, treat_boundary=False, supg=supg)
    M1 = O.scalar_mass(om, np.ones(om.shape))
    ref = (c @ a) * M1
    inner = _interior(om)
    sel = inner if supg else np.ones(om.shape, dtype=bool)  # SUPG adds boundary fluxes at boundary nodes
    assert np.abs(C[sel] - ref[sel]).max() < 1e-12 * np.abs(ref).max()
    C1 = O.supg_apply(om, vel, np.ones(om.shape), kappa=0.0, dt=1.0, mass_scaling=0.0, treat_boundary=False,
                      supg=supg)
    assert np.abs(C1).max() < 1e-12 * np.abs(ref).max()


def test_treated_boundary_rows_and_diagonal_by_brute_force():
    om = O.Mesh(1, 1)
    rng = np.random.default_rng(3)
    vel = consistent(om, rng.uniform(-1, 1, (3,) + om.shape))
    kw = dict(kappa=0.3, dt=0.05, mass_scaling=1.0)
    n = om.nalloc
    # brute-force diagonal of the (untreated, exchanged) operator from unit vectors of unique nodes
    cid = om.canonical_ids.ravel()
    diag = np.zeros(n)
    for u in range(om.nunique):
        e = (cid == u).astype(np.float64).reshape(om.shape)
        out = O.supg_apply(om, vel, e, treat_boundary=False, **kw).ravel()
        diag[cid == u] = out[cid == u]
    d = O.supg_apply(om, vel, np.ones(om.shape), treat_boundary=False, diagonal=True, **kw).ravel()
    assert np.abs(d - diag).max() < 1e-13 * np.abs(diag).max()
    # treated boundary: boundary rows = (their diagonal) x value
    T = consistent(om, rng.uniform(-1, 1, om.shape))
    tb = O.supg_apply(om, vel, T, treat_boundary=True, **kw).ravel()
    dtb = O.supg_apply(om, vel, np.ones(om.shape), treat_boundary=True, diagonal=True, **kw).ravel()
    bnd = om.boundary_mask.ravel() != 1
    assert np.abs(tb[bnd] - dtb[bnd] * T.ravel()[bnd]).max() < 1e-13 * np.abs(tb).max()
    # interior rows with boundary values removed equal the untreated operator on the zeroed field
    Tz = np.where(om.boundary_mask != 1, 0.0, T)
    un = O.supg_apply(om, vel, Tz, treat_boundary=False, **kw).ravel()
    assert np.abs(tb[~bnd] - un[~bnd]).max() < 1e-13 * np.abs(tb).max()
