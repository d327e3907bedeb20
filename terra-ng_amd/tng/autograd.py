"""torch.autograd bridge of the viscosity adjoint (a thin harness over the C ABI).

``StokesVelocity.apply(log_eta, problem)`` returns the velocity u of the FP64
Stokes solve K(exp(log_eta)) [u; p] = f of ``problem``; its backward returns
dL/dlog_eta through ``tng_stokes_adjoint_gradient``.  The forward solution is
kept on the device (checkpoint); Jacobians are recomputed by the kernels.

Tensors hold unique nodes of the velocity mesh: log_eta has shape (n,), u has
shape (3, n); ``UniqueMap`` converts between unique nodes and the library's
node copies (duplicated interface nodes).  torch is used on the host only.
"""
from __future__ import annotations

import numpy as np
import torch

import tng
from tng import inputs


class UniqueMap:
    """Node copies <-> unique nodes of one mesh."""

    def __init__(self, domain):
        info = domain.info
        ids = inputs.node_ids(domain.coords(), info.lateral_level, info.radial_level, info.lateral_sdr,
                              info.radial_sdr, domain.nr)
        uniq, inv = np.unique(ids.ravel(), return_inverse=True)
        self.inv = inv.ravel()
        self.n = uniq.size
        self.rep = np.empty(self.n, dtype=np.int64)
        self.rep[self.inv] = np.arange(self.inv.size)
        self.shape = domain.shape

    def to_copies(self, v):
        v = np.asarray(v)
        return np.ascontiguousarray(v[..., self.inv].reshape(v.shape[:-1] + self.shape))

    def to_unique(self, f):
        f = np.asarray(f)
        return np.ascontiguousarray(f.reshape(f.shape[:-4] + (-1,))[..., self.rep])


class StokesProblem:
    """A Stokes solver with a fixed right-hand side (f_u, f_p given on node copies)."""

    def __init__(self, solver, f_u, f_p=None):
        self.solver = solver
        dv, dp = solver.domain_velocity, solver.domain_pressure
        self.map = UniqueMap(dv)
        self.f_u = tng.VectorQ1Vec(dv).upload(f_u)
        self.f_p = tng.VectorQ1Scalar(dp)
        if f_p is not None:
            self.f_p.upload(f_p)
        self.u = tng.VectorQ1Vec(dv)
        self.p = tng.VectorQ1Scalar(dp)
        self.eta = tng.VectorQ1Scalar(dv)
        self.dJdu = tng.VectorQ1Vec(dv)
        self.g = tng.VectorQ1Scalar(dv)
        self.last_report = None

    def solve(self, eta_unique):
        self.eta.upload(self.map.to_copies(eta_unique))
        self.solver.set_viscosity(self.eta)
        self.u.fill(0.0)
        self.p.fill(0.0)
        self.last_report = self.solver.solve(self.u, self.p, self.f_u, self.f_p)
        return self.map.to_unique(self.u.download())

    def gradient(self, grad_u_unique):
        self.dJdu.upload(self.map.to_copies(grad_u_unique))
        self.last_report = self.solver.adjoint_gradient(self.u, self.dJdu, self.g)
        return self.map.to_unique(self.g.download())


class StokesVelocity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_eta, problem):
        eta = np.exp(log_eta.detach().cpu().numpy().astype(np.float64))
        u = problem.solve(eta)
        ctx.problem = problem
        return torch.from_numpy(u)

    @staticmethod
    def backward(ctx, grad_u):
        g = ctx.problem.gradient(grad_u.detach().cpu().numpy().astype(np.float64))
        return torch.from_numpy(g), None
