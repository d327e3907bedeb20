"""Viscosity adjoint on the GPU (no reference counterpart).

- The coefficient-gradient kernel against the oracle's restatement (pinned by
  homogeneity and finite differences in tests/test_oracle_adjoint_pins.py).
- dJ/dlog(eta) of J = 1/2 ||u - u_obs||^2 (OWNED velocity dofs) through the
  FP64 Stokes solve against central finite differences (h = 1e-6 in log eta,
  solves converged to 1e-13): |g.d - FD| / |FD| <= 1e-6, at level 4 (3
  directions) and at level 5 with eta_true log-normal and 20 directions.
"""
import numpy as np
import pytest

import oracle as O
import tng
from conftest import consistent, rel_err

pytestmark = pytest.mark.gpu

N, D, F = tng.NEUMANN, tng.DIRICHLET, tng.FREESLIP


def field(om, ncomp, seed):
    r = np.random.default_rng(seed)
    return consistent(om, r.uniform(-1, 1, ((ncomp,) if ncomp > 1 else ()) + om.shape))


def vec(dom, ncomp, data=None):
    v = (tng.VectorQ1Vec if ncomp == 3 else tng.VectorQ1Scalar)(dom)
    if data is not None:
        v.upload(data)
    return v


@pytest.mark.parametrize("bcs", [(N, N), (D, D), (F, F), (D, F)])
@pytest.mark.parametrize("L,ls,rs", [(3, 0, 0), (3, 1, 1), (4, 0, 0)])
def test_coefficient_gradient_matches_oracle(bcs, L, ls, rs, gpu_ctx):
    dom = tng.DistributedDomain.create_uniform(L, L, 0.55, 1.0, ls, rs)
    om = O.Mesh(L, L, 0.55, 1.0, ls, rs)
    k = np.exp(field(om, 1, 1))
    lam, u = field(om, 3, 2), field(om, 3, 3)
    kv = vec(dom, 1, k)
    op = tng.EpsilonDivDivKerngen(dom, kv, bcs)
    g = vec(dom, 1)
    op.log_coefficient_gradient(vec(dom, 3, lam), vec(dom, 3, u), g)
    assert rel_err(g.download(), om.log_k_gradient(k, lam, u, *bcs)) < 1e-12


def test_stokes_adjoint_finite_differences(gpu_ctx):
    """Velocity level 4 / pressure level 3, Dirichlet on both surfaces."""
    of = O.Mesh(4)
    own = of.ownership_mask == 1
    S = tng.StokesSolver(2, 4, 0.55, 1.0, (D, D), tng.StokesParams.defaults(
        krylov_restart=50, krylov_max_iterations=600, krylov_relative_tolerance=1e-13,
        krylov_absolute_tolerance=1e-30))
    dv, dp = S.domain_velocity, S.domain_pressure
    z, z2 = field(of, 1, 4), field(of, 1, 5)
    f = field(of, 3, 6)
    fu, fp = vec(dv, 3, f), vec(dp, 1)
    S.enforce_rhs(fu)

    def solve(log_eta):
        S.set_viscosity(vec(dv, 1, np.exp(log_eta)))
        u, p = vec(dv, 3), vec(dp, 1)
        rep = S.solve(u, p, fu, fp)
        assert rep["converged"] == 1, rep
        return u

    u_obs = solve(z).download()
    log_eta = z + 0.1 * z2

    def J(le):
        d = solve(le).download() - u_obs
        return 0.5 * float(np.sum(d[:, own] ** 2))

    u = solve(log_eta)
    dJdu = vec(dv, 3, u.download() - u_obs)  # consistent: each copy holds its node's derivative
    g = vec(dv, 1)
    rep = S.adjoint_gradient(u, dJdu, g)
    assert rep["converged"] == 1, rep
    gh = g.download()
    h = 1e-6
    for seed in (7, 8, 9):
        d = consistent(of, np.random.default_rng(seed).normal(size=of.shape))
        fd = (J(log_eta + h * d) - J(log_eta - h * d)) / (2 * h)
        gd = float(np.sum((gh * d)[own]))
        assert abs(gd - fd) <= 1e-6 * abs(fd), (seed, gd, fd)


def test_adjoint_level5_twenty_directions(gpu_ctx):
    """The adjoint case at its stated size: velocity level 5 / pressure
    level 4, Dirichlet on both surfaces. u_obs is the FP64 solve at
    eta_true = exp(N(0, 1)); the gradient of J with respect to nodal log eta
    is taken at eta_true * exp(0.1 N(0, 1)) and checked against central finite
    differences (h = 1e-6) along 20 seeded directions."""
    of = O.Mesh(5)
    own = of.ownership_mask == 1
    S = tng.StokesSolver(2, 5, 0.55, 1.0, (D, D), tng.StokesParams.defaults(
        krylov_restart=50, krylov_max_iterations=800, krylov_relative_tolerance=1e-13,
        krylov_absolute_tolerance=1e-30))
    dv, dp = S.domain_velocity, S.domain_pressure
    rng = np.random.default_rng(50)
    log_eta_true = consistent(of, rng.normal(size=of.shape))
    log_eta = log_eta_true + 0.1 * consistent(of, rng.normal(size=of.shape))
    fu, fp = vec(dv, 3, field(of, 3, 51)), vec(dp, 1)
    S.enforce_rhs(fu)

    def solve(le):
        S.set_viscosity(vec(dv, 1, np.exp(le)))
        u, p = vec(dv, 3), vec(dp, 1)
        rep = S.solve(u, p, fu, fp)
        assert rep["converged"] == 1, rep
        return u

    u_obs = solve(log_eta_true).download()

    def J(le):
        d = solve(le).download() - u_obs
        return 0.5 * float(np.sum(d[:, own] ** 2))

    u = solve(log_eta)
    g = vec(dv, 1)
    rep = S.adjoint_gradient(u, vec(dv, 3, u.download() - u_obs), g)
    assert rep["converged"] == 1, rep
    gh = g.download()
    h = 1e-6
    for seed in range(100, 120):
        d = consistent(of, np.random.default_rng(seed).normal(size=of.shape))
        fd = (J(log_eta + h * d) - J(log_eta - h * d)) / (2 * h)
        gd = float(np.sum((gh * d)[own]))
        assert abs(gd - fd) <= 1e-6 * abs(fd), (seed, gd, fd)
