"""Pins of the oracle's coefficient gradient d(lam^T A(k) u)/d log k (the
element-wise kernel of the viscosity adjoint).  No GPU needed.

- Homogeneity: A(c k) = c A(k), so the sum of the gradient over the unique
  nodes equals lam^T A(k) u.
- Directional derivatives against central finite differences of the oracle's
  own apply (the slow path, epsilon_divdiv_kerngen.hpp:1013-1242) in log k.
"""
import numpy as np
import pytest

import oracle as O
from conftest import consistent, rel_err

N, D, F = O.NEUMANN, O.DIRICHLET, O.FREESLIP


def field(om, ncomp, seed, scale=1.0):
    r = np.random.default_rng(seed)
    return consistent(om, scale * r.uniform(-1, 1, ((ncomp,) if ncomp > 1 else ()) + om.shape))


def form(m, k, lam, u, bcs):
    return m.masked_dot(lam, m.epsdivdiv(k, u, *bcs, penalty_eps=0.0, path="slow"))


@pytest.mark.parametrize("bcs", [(N, N), (D, D), (F, F), (D, F), (F, N)])
@pytest.mark.parametrize("ls,rs", [(0, 0), (1, 1)])
def test_gradient_homogeneity(bcs, ls, rs):
    m = O.Mesh(2, 2, 0.55, 1.0, ls, rs)
    k = np.exp(field(m, 1, 1))
    lam, u = field(m, 3, 2), field(m, 3, 3)
    g = m.log_k_gradient(k, lam, u, *bcs)
    own = m.ownership_mask == 1
    assert np.abs(g - consistent(m, g)).max() < 1e-14 * np.abs(g).max()
    total = g[own].sum()
    assert abs(total - form(m, k, lam, u, bcs)) < 1e-12 * abs(total)


@pytest.mark.parametrize("bcs", [(D, D), (F, F), (N, F)])
def test_gradient_finite_differences(bcs):
    m = O.Mesh(2)
    k = np.exp(field(m, 1, 4))
    lam, u = field(m, 3, 5), field(m, 3, 6)
    g = m.log_k_gradient(k, lam, u, *bcs)
    own = m.ownership_mask == 1
    for seed in (7, 8):
        d = field(m, 1, seed)
        h = 1e-4
        fd = (form(m, k * np.exp(h * d), lam, u, bcs) - form(m, k * np.exp(-h * d), lam, u, bcs)) / (2 * h)
        gd = (g * d)[own].sum()
        assert abs(gd - fd) < 1e-7 * abs(fd)
