"""torch.autograd harness of the viscosity adjoint (tng/autograd.py): the
gradient of a scalar loss through StokesVelocity matches finite differences.
Collected after the other GPU tests (it imports torch, host side only)."""
import numpy as np
import pytest

import tng

pytestmark = pytest.mark.gpu


def test_autograd_matches_finite_differences(gpu_ctx):
    import torch

    from tng.autograd import StokesProblem, StokesVelocity

    S = tng.StokesSolver(1, 3, 0.55, 1.0, (tng.DIRICHLET, tng.FREESLIP), tng.StokesParams.defaults(
        krylov_restart=50, krylov_max_iterations=600, krylov_relative_tolerance=1e-13,
        krylov_absolute_tolerance=1e-30))
    rng = np.random.default_rng(3)
    prob = StokesProblem(S, np.zeros((3,) + S.domain_velocity.shape))
    n = prob.map.n
    f_unique = rng.uniform(-1, 1, (3, n))
    prob.f_u.upload(prob.map.to_copies(f_unique))
    S.enforce_rhs(prob.f_u)
    w = torch.from_numpy(rng.uniform(-1, 1, (3, n)))
    log_eta = torch.from_numpy(0.3 * rng.normal(size=n)).requires_grad_(True)

    def loss(le):
        return (StokesVelocity.apply(le, prob) * w).sum()

    L = loss(log_eta)
    L.backward()
    g = log_eta.grad.numpy().copy()
    h = 1e-6
    for seed in (4, 5):
        d = torch.from_numpy(np.random.default_rng(seed).normal(size=n))
        with torch.no_grad():
            fd = (loss(log_eta + h * d).item() - loss(log_eta - h * d).item()) / (2 * h)
        assert abs(float(g @ d.numpy()) - fd) <= 1e-6 * abs(fd)
