"""The largest single-GPU case at its stated size: velocity level 7 / pressure
level 6, free slip on both surfaces, viscosity of contrast about 1e6 (a 1e2
radial jump, 1e2 low-viscosity plume columns, log-normal(1) noise), buoyancy
RHS, FGMRES(30) with the FP32 or the FP64 V-cycle, then one adjoint
viscosity-gradient evaluation from the kept forward solution.

With re-discretised coarse operators this case does not reach 1e-8 in a
test's time (1.3e-6 after 1000 iterations, 24 s). The checks are therefore
the ones a fixed budget of 150 iterations allows:
- both solves reduce the FP64 residual by at least 1e3;
- the FP32 V-cycle's residual history stays within a factor 2 of the FP64
  V-cycle's at every iteration (mixed precision costs no convergence);
- the FP32 velocity agrees with the FP64 one to 1e-3 relative, the size of
  the residual both reached;
- the adjoint solve on the same budget reduces its residual by 1e3 and gives
  a finite, nonzero gradient.
"""
import numpy as np
import pytest

import tng
from tng import inputs

pytestmark = pytest.mark.gpu

F = tng.FREESLIP
BUDGET = 150


def test_level7_freeslip_plumes_fixed_budget(gpu_ctx):
    sols, hist, eta, T = {}, {}, None, None
    for name, dt in (("fp64", tng.F64), ("fp32", tng.F32)):
        prm = tng.StokesParams.defaults(krylov_restart=30, krylov_max_iterations=BUDGET,
                                        krylov_relative_tolerance=1e-8, krylov_absolute_tolerance=1e-30,
                                        viscous_pc_chebyshev_order=3, viscous_pc_dtype=dt)
        S = tng.StokesSolver(2, 7, 0.55, 1.0, (F, F), prm)
        dv, dp = S.domain_velocity, S.domain_pressure
        if eta is None:
            w = inputs.Workload(dv.coords(), dv.radii(), 7, 7)
            eta = w.eta_plumes(105)
            T = w.temperature_harmonics(104, 8)
            del w
            print(f"eta range {eta.min():.3e} .. {eta.max():.3e}", flush=True)
        S.set_viscosity(tng.VectorQ1Scalar(dv).upload(eta))
        fu = tng.VectorQ1Vec(dv)
        S.buoyancy_rhs(tng.VectorQ1Scalar(dv).upload(T), 1e5, fu)
        u, p = tng.VectorQ1Vec(dv), tng.VectorQ1Scalar(dp)
        rep = S.solve(u, p, fu, tng.VectorQ1Scalar(dp))
        print(f"{name}-MG: {rep['iterations']} iterations, residual {rep['relative_residual']:.3e}", flush=True)
        assert rep["iterations"] == BUDGET and rep["relative_residual"] <= 1e-3, rep
        hist[name] = np.array(rep["history"])
        own = dv.ownership_mask() == 1
        sols[name] = u.download()[:, own]
        if name == "fp32":
            g = tng.VectorQ1Scalar(dv)
            arep = S.adjoint_gradient(u, u, g)  # J = 1/2 |u|^2, dJ/du = u
            print(f"adjoint: {arep['iterations']} iterations, residual {arep['relative_residual']:.3e}", flush=True)
            assert arep["relative_residual"] <= 1e-3, arep
            gh = g.download()
            assert np.all(np.isfinite(gh)) and np.linalg.norm(gh) > 0
        del S, u, p, fu
    n = min(len(hist["fp32"]), len(hist["fp64"]))
    assert n >= BUDGET
    ratio = hist["fp32"][:n] / hist["fp64"][:n]
    print(f"FP32/FP64 residual ratio over {n} iterations: {ratio.min():.3f} .. {ratio.max():.3f}", flush=True)
    assert np.all(ratio < 2.0) and np.all(ratio > 0.5)
    err = np.linalg.norm(sols["fp32"] - sols["fp64"]) / np.linalg.norm(sols["fp64"])
    print(f"FP32-MG velocity vs FP64-MG: {err:.3e}", flush=True)
    assert err < 1e-3
