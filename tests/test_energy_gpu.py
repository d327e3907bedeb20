"""The caller around the Stokes solve on the GPU through the C ABI:
- the SUPG advection-diffusion operator (unsteady_advection_diffusion_supg.hpp)
  and the scalar mass (mass.hpp) against the oracle (1e-12 relative);
- the SUPG energy step (apps/mantlecirculation/src/energy_solver.hpp:85-276):
  the new temperature solves the eliminated system of the oracle's operators
  to the solver's tolerance and takes the Dirichlet values on the boundary;
  compute_dt = dt_scaling * h / max |u|;
- the time-loop step (mantlecirculation.cpp:413-445): identical bits to the
  same calls made one by one, and a few steps of a small convection case."""
import numpy as np
import pytest

import oracle as O
import tng
from conftest import consistent, rel_err

pytestmark = pytest.mark.gpu

D, F = tng.DIRICHLET, tng.FREESLIP


def field(om, ncomp, seed, scale=1.0):
    r = np.random.default_rng(seed)
    return consistent(om, scale * r.uniform(-1, 1, ((ncomp,) if ncomp > 1 else ()) + om.shape))


def vec(dom, ncomp, data=None):
    v = (tng.VectorQ1Vec if ncomp == 3 else tng.VectorQ1Scalar)(dom)
    if data is not None:
        v.upload(data)
    return v


@pytest.mark.parametrize("L,Lr,ls,rs", [(3, 3, 0, 0), (2, 4, 1, 0), (2, 3, 0, 1), (1, 1, 0, 0)])
@pytest.mark.parametrize("flags", [dict(), dict(treat_boundary=False), dict(diagonal=True),
                                   dict(treat_boundary=False, diagonal=True), dict(lumped_mass=True),
                                   dict(supg=False), dict(mass_scaling=1.5, lumped_mass=True, diagonal=True)])
def test_supg_apply_matches_oracle(L, Lr, ls, rs, flags, gpu_ctx):
    dom = tng.DistributedDomain.create_uniform(L, Lr, 0.55, 1.0, ls, rs)
    om = O.Mesh(L, Lr, 0.55, 1.0, ls, rs)
    vel = field(om, 3, 1, 50.0)
    T = field(om, 1, 2)
    kw = dict(treat_boundary=True, diagonal=False, mass_scaling=1.0, lumped_mass=False)
    supg = flags.pop("supg", True)
    kw.update(flags)
    v = vec(dom, 3, vel)
    op = tng.UnsteadyAdvectionDiffusionSUPG(dom, v, 0.01, 0.02, **kw)
    op.set_supg_enabled(supg)
    out = vec(dom, 1)
    op.apply(vec(dom, 1, T), out)
    ref = O.supg_apply(om, vel, T, kappa=0.01, dt=0.02, mass_scaling=kw["mass_scaling"],
                       treat_boundary=kw["treat_boundary"], diagonal=kw["diagonal"], lumped=kw["lumped_mass"],
                       supg=supg)
    assert rel_err(out.download(), ref) < 1e-12
    # Add mode and skipped communication
    base = field(om, 1, 3)
    op.set_operator_apply_and_communication_modes(tng.ADD, tng.SKIP_COMMUNICATION)
    out2 = vec(dom, 1, base)
    op.apply(vec(dom, 1, T), out2)
    ref2 = O.supg_apply(om, vel, T, kappa=0.01, dt=0.02, mass_scaling=kw["mass_scaling"],
                        treat_boundary=kw["treat_boundary"], diagonal=kw["diagonal"], lumped=kw["lumped_mass"],
                        supg=supg, apply_mode=O.ADD, comm=O.SKIP, dst=base)
    assert rel_err(out2.download(), ref2) < 1e-12


def test_supg_set_dt(gpu_ctx):
    """set_dt (unsteady_advection_diffusion_supg.hpp:206-207): the next apply
    uses the new time step in the mass/advection-diffusion weighting and the
    SUPG parameter."""
    dom = tng.DistributedDomain.create_uniform(3, 3, 0.55, 1.0)
    om = O.Mesh(3)
    vel = field(om, 3, 24, 50.0)
    T = field(om, 1, 25)
    op = tng.UnsteadyAdvectionDiffusionSUPG(dom, vec(dom, 3, vel), 0.01, 0.02)
    out = vec(dom, 1)
    for dt in (0.05, 0.003):
        op.set_dt(dt)
        op.apply(vec(dom, 1, T), out)
        ref = O.supg_apply(om, vel, T, kappa=0.01, dt=dt, mass_scaling=1.0, treat_boundary=True, diagonal=False,
                           lumped=False, supg=True)
        assert rel_err(out.download(), ref) < 1e-12


def test_scalar_mass(gpu_ctx):
    dom = tng.DistributedDomain.create_uniform(3, 3, 0.55, 1.0, 1, 0)
    om = O.Mesh(3, 3, 0.55, 1.0, 1, 0)
    x = field(om, 1, 4)
    out = vec(dom, 1)
    tng.mass_apply(vec(dom, 1, x), out)
    assert rel_err(out.download(), O.scalar_mass(om, x)) < 1e-12


def _convection_setup(L=3, bcs=(D, D), seed=5, speed=200.0):
    S = tng.StokesSolver(2, L, 0.55, 1.0, bcs, tng.StokesParams.defaults(
        krylov_restart=30, krylov_max_iterations=300, krylov_relative_tolerance=1e-8,
        krylov_absolute_tolerance=1e-30))
    dv = S.domain_velocity
    om = O.Mesh(L)
    r = np.broadcast_to(om.radii[:, None, None, :], om.shape)
    # conductive profile plus a perturbation
    T0 = (0.55 * 1.0 / r - 0.55) / 0.45 + 0.05 * field(om, 1, seed)
    T0 = np.clip(T0, 0.0, 1.0)
    u0 = field(om, 3, seed + 1, speed)
    return S, dv, om, T0, u0


def test_energy_step_solves_the_eliminated_system(gpu_ctx):
    S, dv, om, T0, u0 = _convection_setup()
    u = vec(dv, 3, u0)
    T = vec(dv, 1, T0)
    prm = tng.EnergyParams.defaults(krylov_relative_tolerance=1e-10, krylov_max_iterations=500,
                                    krylov_absolute_tolerance=1e-30)
    E = tng.SUPGEnergySolver(S, u, T, prm)
    dt = E.compute_dt()
    h = np.diff(om.radii, axis=1).min()
    assert abs(dt - 0.5 * h / np.linalg.norm(u0, axis=0).max()) < 1e-14 * dt
    rep = E.step(dt)
    assert rep["converged"] == 1, rep
    Tn = T.download()
    # the eliminated system with the oracle's operators
    g = np.where(om.boundary_mask == (2 | 4), 1.0, np.where(om.boundary_mask == (2 | 8), 0.0, 0.0))
    q = O.scalar_mass(om, T0)
    q = q - O.supg_apply(om, u0, g, kappa=1.0, dt=dt, treat_boundary=False)
    gd = O.supg_apply(om, u0, g, kappa=1.0, dt=dt, treat_boundary=False, diagonal=True)
    bnd = om.boundary_mask != 1
    q[bnd] = gd[bnd]
    res = O.supg_apply(om, u0, Tn, kappa=1.0, dt=dt, treat_boundary=True) - q
    own = om.ownership_mask == 1
    assert np.linalg.norm(res[own]) < 1e-9 * np.linalg.norm(q[own])
    # Dirichlet values taken strongly
    assert np.abs(Tn[om.boundary_mask == (2 | 4)] - 1.0).max() < 1e-9
    assert np.abs(Tn[om.boundary_mask == (2 | 8)]).max() < 1e-9


def test_timestep_equals_the_calls_one_by_one(gpu_ctx):
    out = {}
    for mode in ("loop", "manual"):
        S, dv, om, T0, u0 = _convection_setup(seed=7)
        u, p, T = vec(dv, 3, u0), vec(S.domain_pressure, 1), vec(dv, 1, T0)
        E = tng.SUPGEnergySolver(S, u, T)
        if mode == "loop":
            dt, srep, erep = tng.timestep(S, E, u, p, picard_iterations=2, viscosity_law=1, rmu=10.0,
                                          rayleigh_number=1e4)
        else:
            dt = E.compute_dt()
            for it in range(2):
                if it > 0:
                    T.upload(T0)
                fu = vec(dv, 3)
                S.buoyancy_rhs(T, 1e4, fu)
                S.solve(u, p, fu, vec(S.domain_pressure, 1))
                E.step(dt)
                tng.update_viscosity(S, T, 10.0)
        out[mode] = (dt, u.download(), p.download(), T.download())
    for a, b in zip(out["loop"], out["manual"]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("rmu,refresh", [(4.0, False), (100.0, True)])
def test_convection_steps(rmu, refresh, gpu_ctx):
    """Five time steps of a small convection case (no-slip, from rest) with a
    temperature-dependent viscosity: every Stokes and energy solve converges,
    the temperature stays within the boundary values up to the SUPG
    undershoots, and the boundary values hold.  With the reference's update
    (velocity level only) the contrast is kept small (eta in [1/2, 2]); the
    factor-100 contrast runs with the hierarchy rebuilt after each update."""
    S, dv, om, T0, u0 = _convection_setup(seed=9, speed=0.0)
    u, p, T = vec(dv, 3), vec(S.domain_pressure, 1), vec(dv, 1, T0)
    E = tng.SUPGEnergySolver(S, u, T)
    for step in range(5):
        dt, srep, erep = tng.timestep(S, E, u, p, viscosity_law=1, rmu=rmu, rayleigh_number=1e4,
                                      refresh_hierarchy=refresh)
        assert srep["converged"] == 1 and erep["converged"] == 1, (step, srep, erep)
        Tn = T.download()
        assert np.isfinite(Tn).all() and Tn.min() > -0.1 and Tn.max() < 1.1, (step, Tn.min(), Tn.max())
        assert np.abs(Tn[om.boundary_mask == (2 | 4)] - 1.0).max() < 1e-6
        assert dt > 0
