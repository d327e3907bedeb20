"""Pins of the oracle's stored-matrix slow path and TwoGridGCA
(linalg/solvers/gca/gca.hpp:41-531) against closed forms.  No GPU needed.

- Stored Neumann local matrices reproduce the on-the-fly slow path exactly
  (the reference applies stored matrices with the same code,
  epsilon_divdiv_kerngen.hpp:1013-1242).
- The element-wise Galerkin product with exact interpolation equals the
  global product R A_fine P on consistent vectors (R = P^T), over one and
  two coarsening steps.
- The reference's own rule (gca.hpp:206-208) places the weights of fine
  nodes on the top face of an interior coarse cell on the coarse cell's
  bottom nodes.  It coincides with exact interpolation wherever that face is
  the shell's surface, so it matches exactly on single-layer shells and on
  the top layer of cells, and differs below it.
- The adaptive-GCA collector (gca_elements_collector.hpp:20-96) on the
  polynomial field k = a x^2 + b y^2 + c r^2 + d x r over the node indices
  flags exactly the cells a closed-form centroid gradient puts above 10,
  mapped to the ancestor cell; the selective assembly takes the Galerkin
  matrices exactly on flagged ancestors.
"""
import numpy as np
import pytest

import oracle as O
from conftest import consistent, rel_err

N, D, F = O.NEUMANN, O.DIRICHLET, O.FREESLIP


def field(om, ncomp, seed):
    r = np.random.default_rng(seed)
    return consistent(om, r.uniform(-1, 1, ((ncomp,) if ncomp > 1 else ()) + om.shape))


@pytest.mark.parametrize("bcs", [(N, N), (D, D), (F, F), (D, F), (F, N)])
@pytest.mark.parametrize("diagonal", [False, True])
def test_stored_local_matrices_match_slow_path(bcs, diagonal):
    m = O.Mesh(2, 2, 0.55, 1.0, 1, 0)
    k = np.exp(field(m, 1, 1))
    u = field(m, 3, 2)
    lm = O.local_matrices(m, k)
    a = O.epsdivdiv_stored(m, lm, u, *bcs, diagonal=diagonal)
    b = m.epsdivdiv(k, u, *bcs, diagonal=diagonal, path="slow")
    assert np.array_equal(a, b)


def galerkin(mf, mc, lm_f, x):
    """R A_fine P x with the Neumann stored fine matrices."""
    px = O.prolongate(mc, mf, x)
    return O.restrict(mf, mc, O.epsdivdiv_stored(mf, lm_f, px, N, N))


@pytest.mark.parametrize("lat,rad,ls", [(2, 1, 0), (3, 1, 0), (2, 2, 0), (3, 2, 1)])
def test_gca_exact_is_global_galerkin_product(lat, rad, ls):
    mf = O.Mesh(lat, rad, 0.55, 1.0, ls, 0)
    mc = O.Mesh(lat - 1, rad - 1, 0.55, 1.0, ls, 0)
    k = np.exp(field(mf, 1, 3))
    lm_f = O.local_matrices(mf, k)
    lm_c = O.gca(mf, mc, lm_f, exact_top=True)
    for seed in (4, 5):
        x = field(mc, 3, seed)
        got = O.epsdivdiv_stored(mc, lm_c, x, N, N)
        assert rel_err(got, galerkin(mf, mc, lm_f, x)) < 1e-13
    # the coarse local matrices are symmetric like the fine ones
    assert np.abs(lm_c - np.swapaxes(lm_c, -1, -2)).max() < 1e-13 * np.abs(lm_c).max()


def test_gca_two_steps_exact():
    m2, m1, m0 = O.Mesh(3, 2), O.Mesh(2, 1), O.Mesh(1, 0)
    k = np.exp(field(m2, 1, 6))
    l2 = O.local_matrices(m2, k)
    l1 = O.gca(m2, m1, l2, exact_top=True)
    l0 = O.gca(m1, m0, l1, exact_top=True)
    x = field(m0, 3, 7)
    ref = O.restrict(m1, m0, O.restrict(m2, m1, O.epsdivdiv_stored(
        m2, l2, O.prolongate(m1, m2, O.prolongate(m0, m1, x)), N, N)))
    assert rel_err(O.epsdivdiv_stored(m0, l0, x, N, N), ref) < 1e-13


@pytest.mark.parametrize("lat,rad", [(2, 1), (3, 1)])
def test_gca_reference_rule_single_layer(lat, rad):
    """On a single-layer coarse shell the reference rule is exact."""
    mf, mc = O.Mesh(lat, rad), O.Mesh(lat - 1, rad - 1)
    lm_f = O.local_matrices(mf, np.exp(field(mf, 1, 8)))
    ref = O.gca(mf, mc, lm_f)
    assert np.array_equal(ref, O.gca(mf, mc, lm_f, exact_top=True))


def test_gca_reference_rule_differs_below_top_layer():
    mf, mc = O.Mesh(3, 3), O.Mesh(2, 2)
    lm_f = O.local_matrices(mf, np.exp(field(mf, 1, 9)))
    ref = O.gca(mf, mc, lm_f)
    exact = O.gca(mf, mc, lm_f, exact_top=True)
    top = mc.nr - 2
    assert np.array_equal(ref[:, :, :, top], exact[:, :, :, top])
    for rc in range(top):
        assert np.abs(ref[:, :, :, rc] - exact[:, :, :, rc]).max() > 1e-3 * np.abs(exact).max()
    # the same P acts on both sides, so the coarse matrices stay symmetric
    assert np.abs(ref - np.swapaxes(ref, -1, -2)).max() < 1e-13 * np.abs(ref).max()


def closed_form_cell_norms(m, a, b, c, d):
    """Largest |grad_ref k| of the two wedges of every cell for
    k = a x^2 + b y^2 + c r^2 + d x r (x, y, r node indices).  Wedge 0 has
    corners (x, y), (x+1, y), (x, y+1), wedge 1 (x+1, y+1), (x, y+1), (x+1, y);
    at the centroid the interpolant's lateral differences are taken at mid
    radius and the radial one at the lateral centroid (x + 1/3 or x + 2/3)."""
    xc = np.arange(m.nx - 1)[:, None, None]
    yc = np.arange(m.nx - 1)[None, :, None]
    rc = np.arange(m.nr - 1)[None, None, :]
    g0 = a * (2 * xc + 1) + d * (rc + 0.5)
    g1 = b * (2 * yc + 1)
    n0 = np.sqrt(g0 ** 2 + g1 ** 2 + (0.5 * (c * (2 * rc + 1) + d * (xc + 1.0 / 3.0))) ** 2)
    n1 = np.sqrt(g0 ** 2 + g1 ** 2 + (0.5 * (c * (2 * rc + 1) + d * (xc + 2.0 / 3.0))) ** 2)
    return np.broadcast_to(np.maximum(n0, n1), (m.nsd, m.nx - 1, m.nx - 1, m.nr - 1))


def polynomial_k(m, a, b, c, d):
    x = np.arange(m.nx)[:, None, None]
    y = np.arange(m.nx)[None, :, None]
    r = np.arange(m.nr)[None, None, :]
    return np.broadcast_to(a * x ** 2 + b * y ** 2 + c * r ** 2 + d * x * r, m.shape).copy()


@pytest.mark.parametrize("coef", [(0.97, 0.0, 0.0, 0.0), (0.0, 1.07, 0.0, 0.0), (0.0, 0.0, 2.37, 0.0),
                                  (0.0, 0.0, 0.0, 1.93), (0.29, -0.41, 1.13, 0.71)])
@pytest.mark.parametrize("level_range", [0, 1, 2])
def test_gca_elements_closed_form(coef, level_range):
    m = O.Mesh(3, 3)
    coarsest = O.Mesh(3 - level_range, 3 - level_range)
    got = O.gca_elements(m, polynomial_k(m, *coef), level_range, coarsest)
    hit = closed_form_cell_norms(m, *coef) > 10.0
    want = np.zeros(got.shape, dtype=bool)
    sd, x, y, r = np.nonzero(hit)
    want[sd, x >> level_range, y >> level_range, r >> level_range] = True
    assert 0 < want.sum() < want.size  # the case separates flagged and unflagged cells
    assert np.array_equal(got.astype(bool), want)


def test_gca_selective_takes_galerkin_on_flagged_ancestors():
    mf, mc, m0 = O.Mesh(3, 3), O.Mesh(2, 2), O.Mesh(1, 1)
    k = np.exp(2.0 * field(mf, 1, 10))
    kc = np.ascontiguousarray(k[:, ::2, ::2, ::2])
    flags = np.zeros((m0.nsd, m0.nx - 1, m0.nx - 1, m0.nr - 1), dtype=np.uint8)
    flags[3, 1, 0, 1] = flags[7, 0, 1, 0] = 1
    sel = O.gca_selective(mf, mc, O.local_matrices(mf, k), kc, 1, flags)
    g, loc = O.gca(mf, mc, O.local_matrices(mf, k)), O.local_matrices(mc, kc)
    for (sd, x0, y0, r0) in np.ndindex(flags.shape):
        blk = np.s_[sd, 2 * x0:2 * x0 + 2, 2 * y0:2 * y0 + 2, 2 * r0:2 * r0 + 2]
        assert np.array_equal(sel[blk], (g if flags[sd, x0, y0, r0] else loc)[blk])
