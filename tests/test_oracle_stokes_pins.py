"""Pins of the oracle's divergence B, gradient B^T, constant prolongation P,
restriction R and lumped k-weighted mass (oracle/tng_oracle.c) against what
the mathematics and the reference's own tests fix.  No GPU needed.

What the reference's tests pin for these operators: constants are
prolongated exactly (tests/test_prolongation_constant.cpp:213-215), R = P^T
(tests/test_restriction.cpp:83-86, 137-140), B and B^T fast == slow
(tests/test_divergence_fast_vs_slow.cpp:231-278,
tests/test_gradient_fast_vs_slow.cpp:210-241).  None asserts output values,
so the pins below add closed forms (discrete volume, exact integrals of the
lumped mass), invariants (B of constants and rotations, transposition) and
an independent brute-force B^T built from the isoparametric map by exact
differences and from geometric barycentric coordinates.

The centre-triangle convention of B / B^T (the centre fine triangle of each
coarse wedge is integrated with the opposite coarse wedge's nodes, weight
1/3 each; divergence_kerngen.hpp:758-775, kernel_helpers.hpp:600-608) is
restated in the brute force from that description; only running the
reference itself would pin it further.
"""
import numpy as np
import pytest

import oracle as O
from conftest import consistent, rel_err

N, D, F = O.NEUMANN, O.DIRICHLET, O.FREESLIP

_meshes = {}


def pair(L, ls=0, rs=0):
    key = (L, ls, rs)
    if key not in _meshes:
        _meshes[key] = (O.Mesh(L, L, 0.55, 1.0, ls, rs), O.Mesh(L - 1, L - 1, 0.55, 1.0, ls, rs))
    return _meshes[key]


def rng_field(mesh, ncomp, seed):
    r = np.random.default_rng(seed)
    shape = ((ncomp,) if ncomp > 1 else ()) + mesh.shape
    return consistent(mesh, r.uniform(-1, 1, shape))


def position(mesh):
    """Physical node positions x = r * unit lateral coordinate, (3, sd, x, y, r)."""
    c, rad = mesh.coords, mesh.radii
    return np.moveaxis(c[:, :, :, None, :] * rad[:, None, None, :, None], -1, 0).copy()


def unit_normal(mesh):
    c = mesh.coords
    return np.broadcast_to(np.moveaxis(c, -1, 0)[..., None], (3,) + mesh.shape)


# ------------------------------------------------------------------ P and R
@pytest.mark.parametrize("ls,rs", [(0, 0), (1, 1)])
def test_prolongation_constants(ls, rs):
    """Constants are prolongated exactly (test_prolongation_constant.cpp:213-215)."""
    mf, mc = pair(3, ls, rs)
    for nc in (1, 3):
        x = np.full(((nc,) if nc > 1 else ()) + mc.shape, 2.5)
        y = O.prolongate(mc, mf, x)
        assert np.abs(y - 2.5).max() < 1e-15


@pytest.mark.parametrize("ls,rs", [(0, 0), (1, 1)])
def test_prolongation_reproduces_mesh(ls, rs):
    """Every fine node is the normalised midpoint of its coarse parents and
    the shell radii are uniform, so P of the coarse unit coordinates points
    along the fine coordinates and P of the coarse radii gives the fine
    radii.  A wrong parent (for example the other diagonal of a cell) fails."""
    mf, mc = pair(3, ls, rs)
    cc = np.broadcast_to(np.moveaxis(mc.coords, -1, 0)[..., None], (3,) + mc.shape).copy()
    pc = O.prolongate(mc, mf, cc)
    pn = pc / np.linalg.norm(pc, axis=0)
    assert np.abs(pn - unit_normal(mf)).max() < 1e-15
    rc = np.broadcast_to(mc.radii[:, None, None, :], mc.shape).copy()
    pr = O.prolongate(mc, mf, rc)
    assert np.abs(pr - np.broadcast_to(mf.radii[:, None, None, :], mf.shape)).max() < 1e-15


@pytest.mark.parametrize("ls,rs", [(0, 0), (1, 0), (1, 1)])
def test_restriction_is_transpose_of_prolongation(ls, rs):
    """<R x, y>_coarse == <x, P y>_fine over OWNED nodes for consistent x, y
    (R = P^T, test_restriction.cpp:83-86)."""
    mf, mc = pair(3, ls, rs)
    for nc, seed in ((1, 11), (3, 12)):
        x = rng_field(mf, nc, seed)
        y = rng_field(mc, nc, seed + 100)
        rx = O.restrict(mf, mc, x)
        py = O.prolongate(mc, mf, y)
        a, b = mc.masked_dot(rx, y), mf.masked_dot(x, py)
        assert abs(a - b) <= 1e-13 * abs(b)
        # the restricted field is consistent across duplicated copies
        assert np.abs(rx - consistent(mc, rx)).max() < 1e-14


def test_prolongation_add_mode():
    mf, mc = pair(3)
    y = rng_field(mc, 3, 5)
    base = rng_field(mf, 3, 6)
    out = O.prolongate(mc, mf, y, apply_mode=O.ADD, dst=base)
    assert np.abs(out - (base + O.prolongate(mc, mf, y))).max() < 1e-15


# ------------------------------------------------------------------ lumped KMass
def _lat_wedges(nx):
    for x in range(nx - 1):
        for y in range(nx - 1):
            yield (x, y), [(x, y), (x + 1, y), (x, y + 1)]
            yield (x, y), [(x + 1, y + 1), (x, y + 1), (x + 1, y)]


@pytest.mark.parametrize("ls,rs", [(0, 0), (1, 1)])
def test_kmass_lumped_exact_integrals(ls, rs):
    """For k varying only laterally, the lumped row sum of node i is the exact
    integral of k N_i over the isoparametric wedges:
        |det[p0,p1,p2]| * sum_b k_b (1 + delta_ab) / 24 * int N_l(r) r^2 dr
    (the 6-point rule is exact for this integrand: quadratic on the triangle,
    cubic in the radius)."""
    m = O.Mesh(2, 2, 0.55, 1.0, ls, rs)
    c, rad = m.coords, m.radii
    klat = 2.0 + np.sin(3 * c[..., 0]) + np.cos(2 * c[..., 1]) * c[..., 2]
    k = np.broadcast_to(klat[..., None], m.shape).copy()
    ones = np.ones(m.shape)
    got = O.kmass_lumped(m, k, ones)
    exp = np.zeros(m.shape)
    for sd in range(m.nsd):
        for _, corners in _lat_wedges(m.nx):
            P = np.array([c[sd, x, y] for x, y in corners])
            det = abs(np.linalg.det(P))
            kb = np.array([klat[sd, x, y] for x, y in corners])
            lat = np.array([(kb.sum() + kb[a]) / 24.0 for a in range(3)])
            for r in range(m.nr - 1):
                r1, r2 = rad[sd, r], rad[sd, r + 1]
                dr = r2 - r1
                i3, i4 = (r2 ** 3 - r1 ** 3) / 3, (r2 ** 4 - r1 ** 4) / 4
                rad_int = ((r2 * i3 - i4) / dr, (i4 - r1 * i3) / dr)  # bottom, top
                for a, (x, y) in enumerate(corners):
                    exp[sd, x, y, r] += det * lat[a] * rad_int[0]
                    exp[sd, x, y, r + 1] += det * lat[a] * rad_int[1]
    exp = m.halo_sum(exp)
    assert rel_err(got, exp) < 1e-14
    # k = 1: the row sums add up to the shell volume of the flat-faceted sphere
    vol = O.kmass_lumped(m, np.ones(m.shape), ones)
    own = m.ownership_mask == 1
    v = 0.0
    for sd in range(m.nsd):
        for _, corners in _lat_wedges(m.nx):
            v += abs(np.linalg.det(np.array([c[sd, x, y] for x, y in corners]))) * (
                rad[sd, -1] ** 3 - rad[sd, 0] ** 3) / 6.0
    assert abs(vol[own].sum() - v) < 1e-14 * v


def test_kmass_scales_src():
    m = O.Mesh(2)
    k = np.exp(rng_field(m, 1, 3))
    x = rng_field(m, 1, 4)
    d = O.kmass_lumped(m, k, np.ones(m.shape))
    assert np.abs(O.kmass_lumped(m, k, x) - d * x).max() < 1e-15 * np.abs(d).max()


# ------------------------------------------------------------------ B and B^T
BC_SETS = [(N, N), (D, D), (F, F), (D, F), (F, N)]


@pytest.mark.parametrize("bcs", BC_SETS)
@pytest.mark.parametrize("ls,rs", [(0, 0), (1, 1)])
def test_gradient_is_transpose_of_divergence(bcs, ls, rs):
    """<B u, p> == <u, B^T p> over OWNED nodes for consistent u, p with the
    same boundary conditions (two separately written restatements)."""
    mf, mc = pair(3, ls, rs)
    u = rng_field(mf, 3, 21)
    p = rng_field(mc, 1, 22)
    bu = O.divergence(mf, mc, u, *bcs)
    gp = O.gradient(mf, mc, p, *bcs)
    a, b = mc.masked_dot(bu, p), mf.masked_dot(u, gp)
    assert abs(a - b) <= 1e-13 * max(abs(a), abs(b)), (a, b)


def test_divergence_kernel_and_volume():
    """Neumann: B annihilates constants and rigid rotations (their nodal
    interpolants are exact and divergence free); for the position field
    div x = 3, so the sum of B x over the unique coarse nodes is -3 times
    the one-point-rule volume sum_wedges r_mid^2 (dr/2) |det[p0,p1,p2]|."""
    mf, mc = pair(3)
    own = mc.ownership_mask == 1
    X = position(mf)
    scale = np.abs(O.divergence(mf, mc, X, N, N)).max()
    const = np.stack([np.full(mf.shape, v) for v in (1.0, -2.0, 0.5)])
    assert np.abs(O.divergence(mf, mc, const, N, N)).max() < 1e-14 * scale
    for e in np.eye(3):
        rot = np.cross(e[:, None], X.reshape(3, -1), axis=0).reshape(X.shape)
        assert np.abs(O.divergence(mf, mc, rot, N, N)).max() < 1e-14 * scale
    c, rad = mf.coords, mf.radii
    vol = 0.0
    for sd in range(mf.nsd):
        dets = sum(abs(np.linalg.det(np.array([c[sd, x, y] for x, y in cr]))) for _, cr in _lat_wedges(mf.nx))
        for r in range(mf.nr - 1):
            rmid, h = 0.5 * (rad[sd, r] + rad[sd, r + 1]), 0.5 * (rad[sd, r + 1] - rad[sd, r])
            vol += rmid * rmid * h * dets
    s = O.divergence(mf, mc, X, N, N)[own].sum()
    assert abs(s + 3 * vol) < 1e-13 * vol


@pytest.mark.parametrize("bcs", [(D, D), (F, F), (D, F), (F, D), (N, F)])
def test_boundary_handling_relations(bcs):
    """Dirichlet removes the boundary velocity from B's trial side and zeroes
    B^T's boundary rows; free slip projects the trial velocity / the B^T
    rows onto the tangent plane (n = unit lateral coordinate)."""
    mf, mc = pair(3)
    u = rng_field(mf, 3, 31)
    p = rng_field(mc, 1, 32)
    n = unit_normal(mf)
    um = u.copy()
    layers = {0: bcs[0], mf.nr - 1: bcs[1]}
    for l, bc in layers.items():
        if bc == D:
            um[..., l] = 0
        elif bc == F:
            d = (um[..., l] * n[..., l]).sum(0)
            um[..., l] -= d * n[..., l]
    assert rel_err(O.divergence(mf, mc, u, *bcs), O.divergence(mf, mc, um, N, N)) < 1e-13
    g = O.gradient(mf, mc, p, N, N)
    for l, bc in layers.items():
        if bc == D:
            g[..., l] = 0
        elif bc == F:
            d = (g[..., l] * n[..., l]).sum(0)
            g[..., l] -= d * n[..., l]
    assert rel_err(O.gradient(mf, mc, p, *bcs), g) < 1e-13


def _brute_force_gradient(mf, mc, p):
    """Neumann B^T p per copy (no exchange) from the isoparametric map by
    exact central differences (the map is linear in each reference variable)
    and coarse shape values from geometric barycentric coordinates."""
    c, rad = mf.coords, mf.radii
    out = np.zeros((3,) + mf.shape)

    def nlat(xi, eta):
        return np.array([1 - xi - eta, xi, eta])

    def nrad(z):
        return np.array([0.5 * (1 - z), 0.5 * (1 + z)])

    def shapes(q):
        return np.outer(nrad(q[2]), nlat(q[0], q[1])).ravel()  # j = 3*e + k

    for sd in range(mf.nsd):
        for (x, y), corners in _lat_wedges(mf.nx):
            w = 0 if corners[0] == (x, y) else 1
            for r in range(mf.nr - 1):
                Xn = np.array([c[sd, a, b] * rad[sd, r + e] for e in (0, 1) for a, b in corners])  # (6, 3)
                q0 = np.array([1 / 3, 1 / 3, 0.0])
                J = np.zeros((3, 3))
                G = np.zeros((6, 3))
                for v in range(3):
                    dq = np.zeros(3)
                    dq[v] = 1.0
                    dN = 0.5 * (shapes(q0 + dq) - shapes(q0 - dq))
                    G[:, v] = dN
                    J[:, v] = dN @ Xn
                grads = G @ np.linalg.inv(J)  # rows: grad N_j (physical)
                detJ = abs(np.linalg.det(J))
                # coarse pressure at the fine centroid
                cxf = np.mean([a for a, _ in corners]) if w == 0 else x + 2 / 3
                cyf = np.mean([b for _, b in corners]) if w == 0 else y + 2 / 3
                X0, Y0, R0 = x // 2, y // 2, r // 2
                s, t = cxf / 2 - X0, cyf / 2 - Y0
                rho = (r + 0.5) / 2 - R0
                centre = (w == 1 and s + t < 1) or (w == 0 and s + t > 1)
                if w == 0:
                    cn = [(X0, Y0), (X0 + 1, Y0), (X0, Y0 + 1)]
                    bary = np.array([1 - s - t, s, t])
                else:
                    cn = [(X0 + 1, Y0 + 1), (X0, Y0 + 1), (X0 + 1, Y0)]
                    bary = np.array([s + t - 1, 1 - s, 1 - t])
                if centre:
                    bary = np.full(3, 1 / 3)
                pc = 0.0
                for e, re in ((0, 1 - rho), (1, rho)):
                    for k3, (a, b) in enumerate(cn):
                        pc += bary[k3] * re * p[sd, a, b, R0 + e]
                for j in range(6):
                    a, b = corners[j % 3]
                    out[:, sd, a, b, r + j // 3] += -detJ * pc * grads[j]
    return out


@pytest.mark.parametrize("ls,rs", [(0, 0), (1, 1)])
def test_gradient_brute_force(ls, rs):
    mf, mc = pair(2, ls, rs)
    p = np.random.default_rng(41).uniform(-1, 1, mc.shape)  # per-copy values, no exchange
    got = O.gradient(mf, mc, p, N, N, comm=O.SKIP)
    exp = _brute_force_gradient(mf, mc, p)
    assert rel_err(got, exp) < 1e-13


def test_divergence_apply_modes():
    mf, mc = pair(3)
    u = rng_field(mf, 3, 51)
    base = rng_field(mc, 1, 52)
    loc = O.divergence(mf, mc, u, D, F, comm=O.SKIP)
    # Add + exchange: the exchange sums the accumulated dst over the copies
    out = O.divergence(mf, mc, u, D, F, apply_mode=O.ADD, comm=O.COMMUNICATE, dst=base)
    assert np.abs(out - mc.halo_sum(base + loc)).max() < 1e-14


# ------------------------------------------------------------------ VectorMass
def test_vector_mass_row_sums_are_lumped_kmass():
    """With k = 1 the row sums of the consistent mass are the lumped mass."""
    m = O.Mesh(2, 2, 0.55, 1.0, 1, 0)
    ones = np.ones((3,) + m.shape)
    lumped = O.kmass_lumped(m, np.ones(m.shape), np.ones(m.shape))
    mm = O.vector_mass(m, ones)
    for d in range(3):
        assert rel_err(mm[d], lumped) < 1e-14


def test_vector_mass_exact_for_lateral_fields():
    """For radially constant fields the 6-point rule is exact:
    (M u)_i = sum_w |det P| sum_b u_b (1 + delta_ab)/24 * int N_l(r) r^2 dr."""
    m = O.Mesh(2)
    c, rad = m.coords, m.radii
    ulat = np.stack([np.sin(2 * c[..., 0]) + c[..., 1], c[..., 2] ** 2, 1.0 + 0.3 * c[..., 0] * c[..., 1]])
    u = np.broadcast_to(ulat[..., None], (3,) + m.shape).copy()
    exp = np.zeros((3,) + m.shape)
    for sd in range(m.nsd):
        for _, corners in _lat_wedges(m.nx):
            det = abs(np.linalg.det(np.array([c[sd, x, y] for x, y in corners])))
            for r in range(m.nr - 1):
                r1, r2 = rad[sd, r], rad[sd, r + 1]
                i3, i4 = (r2 ** 3 - r1 ** 3) / 3, (r2 ** 4 - r1 ** 4) / 4
                rint = ((r2 * i3 - i4) / (r2 - r1), (i4 - r1 * i3) / (r2 - r1))
                for a, (x, y) in enumerate(corners):
                    for d in range(3):
                        ub = np.array([ulat[d, sd, xx, yy] for xx, yy in corners])
                        lat = (ub.sum() + ub[a]) / 24.0
                        for e in (0, 1):
                            # radially constant u: int N_e(r) r^2 dr against the sum of both layers
                            exp[d, sd, x, y, r + e] += det * lat * rint[e]
    exp = m.halo_sum(exp)
    assert rel_err(O.vector_mass(m, u), exp) < 1e-14
