"""Stored local matrices, the stored-matrix slow path and TwoGridGCA on the GPU
against the oracle (epsilon_divdiv_kerngen.hpp:503-575, 1013-1242, 2315-2392;
linalg/solvers/gca/gca.hpp:41-531), adaptive GCA (gca_elements_collector.hpp
:20-96, Selective storage).  FP64 bar 1e-12 relative (matrices 1e-13), FP32
operators 2e-6, element flags bit-exact."""
import numpy as np
import pytest

import oracle as O
import tng
from conftest import consistent, rel_err

pytestmark = pytest.mark.gpu

N, D, F = tng.NEUMANN, tng.DIRICHLET, tng.FREESLIP
_cache = {}


def setup(lat, rad=None, ls=0, rs=0):
    rad = lat if rad is None else rad
    key = (lat, rad, ls, rs)
    if key not in _cache:
        _cache[key] = (tng.DistributedDomain.create_uniform(lat, rad, 0.55, 1.0, ls, rs),
                       O.Mesh(lat, rad, 0.55, 1.0, ls, rs))
    return _cache[key]


def field(om, ncomp, seed):
    r = np.random.default_rng(seed)
    return consistent(om, r.uniform(-1, 1, ((ncomp,) if ncomp > 1 else ()) + om.shape))


def vec(dom, ncomp, data=None, dtype=tng.F64):
    v = (tng.VectorQ1Vec if ncomp == 3 else tng.VectorQ1Scalar)(dom, dtype)
    if data is not None:
        v.upload(data)
    return v


def op_on(dom, k, bcs, dtype=tng.F64):
    kv = vec(dom, 1, k, dtype)
    return tng.EpsilonDivDivKerngen(dom, kv, bcs), kv


@pytest.mark.parametrize("lat,ls,rs", [(2, 0, 0), (3, 1, 1)])
def test_local_matrices_match_oracle(lat, ls, rs, gpu_ctx):
    dom, om = setup(lat, lat, ls, rs)
    k = np.exp(field(om, 1, 1))
    op, kv = op_on(dom, k, (N, N))
    got = op.get_local_matrices()
    ref = O.local_matrices(om, k)
    assert rel_err(got, ref) < 1e-13


@pytest.mark.parametrize("bcs", [(N, N), (D, D), (F, F), (D, F), (F, N)])
@pytest.mark.parametrize("diagonal", [False, True])
def test_stored_apply_matches_slow_path(bcs, diagonal, gpu_ctx):
    dom, om = setup(3, 3, 1, 0)
    k = np.exp(field(om, 1, 2))
    u = field(om, 3, 3)
    lm = O.local_matrices(om, k) * (1.0 + 0.1 * np.random.default_rng(4).uniform(size=(1,) * 5 + (18, 18)))
    op, kv = op_on(dom, k, bcs, tng.F64)
    op.set_local_matrices(lm)
    op.set_diagonal(diagonal)
    out = vec(dom, 3)
    op.apply(vec(dom, 3, u), out)
    ref = O.epsdivdiv_stored(om, lm, u, *bcs, diagonal=diagonal)
    assert rel_err(out.download(), ref) < 1e-12
    # FP32 operator with converted matrices
    op32, kv32 = op_on(dom, k, bcs, tng.F32)
    op32.copy_local_matrices(op)
    op32.set_diagonal(diagonal)
    out32 = vec(dom, 3, dtype=tng.F32)
    op32.apply(vec(dom, 3, u, tng.F32), out32)
    assert rel_err(out32.download(), ref) < 2e-6


def test_stored_apply_modes(gpu_ctx):
    dom, om = setup(3, 3, 0, 1)
    k = np.exp(field(om, 1, 5))
    u, base = field(om, 3, 6), field(om, 3, 7)
    lm = O.local_matrices(om, k)
    op, kv = op_on(dom, k, (F, F))
    op.set_local_matrices(lm)
    op.set_penalty_epsilon(0.5)
    for mode, comm in [(tng.ADD, tng.COMMUNICATE_ADDITIVELY), (tng.REPLACE, tng.SKIP_COMMUNICATION)]:
        op.set_operator_apply_and_communication_modes(mode, comm)
        out = vec(dom, 3, base)
        op.apply(vec(dom, 3, u), out)
        ref = O.epsdivdiv_stored(om, lm, u, F, F, apply_mode=mode, comm=comm, penalty_eps=0.5, dst=base)
        assert rel_err(out.download(), ref) < 1e-12


@pytest.mark.parametrize("lat,rad,ls,rs", [(3, 3, 0, 0), (4, 4, 0, 0), (3, 2, 1, 0), (2, 2, 0, 0), (2, 3, 0, 0),
                                          (3, 3, 0, 1), (4, 3, 1, 1)])
@pytest.mark.parametrize("exact", [False, True])
def test_gca_matches_oracle(lat, rad, ls, rs, exact, gpu_ctx):
    """Both kernels (fine matrices assembled on the fly, then read from
    storage), coarse-cell counts that do and do not fill whole workgroups,
    radially split subdomains (top cells inside the shell), the reference's
    top-face rule and the exact Galerkin product."""
    df, of = setup(lat, rad, ls, rs)
    dc, oc = setup(lat - 1, rad - 1, ls, rs)
    k = np.exp(field(of, 1, 8))
    fine, kf = op_on(df, k, (D, D))
    coarse, kc = op_on(dc, np.ones(oc.shape), (D, D))
    tng.gca_assemble(fine, coarse, exact=exact)  # fine matrices assembled on the fly
    got = coarse.get_local_matrices()
    ref = O.gca(of, oc, O.local_matrices(of, k), exact_top=exact)
    assert rel_err(got, ref) < 1e-13
    # second step from stored matrices (needs two radial layers on the coarser mesh)
    if (2 ** (rad - 2)) >> rs >= 2 and lat - 2 >= 1:
        dcc, occ = setup(lat - 2, rad - 2, ls, rs)
        c2, kc2 = op_on(dcc, np.ones(occ.shape), (D, D))
        tng.gca_assemble(coarse, c2, exact=exact)
        assert rel_err(c2.get_local_matrices(), O.gca(oc, occ, ref, exact_top=exact)) < 1e-13


def test_gca_operator_apply_freeslip(gpu_ctx):
    df, of = setup(4)
    dc, oc = setup(3)
    k = np.exp(field(of, 1, 9))
    fine, kf = op_on(df, k, (F, F))
    coarse, kc = op_on(dc, np.ones(oc.shape), (F, F))
    tng.gca_assemble(fine, coarse)
    x = field(oc, 3, 10)
    out = vec(dc, 3)
    coarse.apply(vec(dc, 3, x), out)
    ref = O.epsdivdiv_stored(oc, O.gca(of, oc, O.local_matrices(of, k)), x, F, F)
    assert rel_err(out.download(), ref) < 1e-12


def contrast_k(om, seed, scale, jump_r=0.8, inner=40.0, outer=1.0):
    """Lognormal viscosity with a radial jump (factor `inner` below jump_r,
    `outer` above): flags some coarsest cells, not all."""
    r = np.broadcast_to(om.radii[:, None, None, :], om.shape)
    return np.where(r < jump_r, inner, outer) * np.exp(scale * field(om, 1, seed))


@pytest.mark.parametrize("lat,level_range,scale", [(3, 1, 1.0), (4, 2, 1.0), (4, 1, 2.0), (2, 0, 3.0)])
def test_gca_elements_collect_matches_oracle(lat, level_range, scale, gpu_ctx):
    df, of = setup(lat)
    d0, o0 = setup(lat - level_range)
    k = contrast_k(of, 13 + lat, scale)
    got = tng.gca_elements_collect(vec(df, 1, k), level_range, d0)
    ref = O.gca_elements(of, k, level_range, o0)
    assert 0 < ref.sum() < ref.size
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("lat,level_range,kargs", [(4, 1, (1.0, 0.8, 40.0, 1.0)), (4, 0, (1.0, 0.8, 40.0, 1.0)),
                                                   (3, 1, (0.3, 0.95, 1.0, 25.0))])
def test_gca_assemble_selective_matches_oracle(lat, level_range, kargs, gpu_ctx):
    df, of = setup(lat)
    dc, oc = setup(lat - 1)
    d0, o0 = setup(lat - 1 - level_range)
    k = contrast_k(of, 20 + lat, *kargs)
    kc = np.ascontiguousarray(k[:, ::2, ::2, ::2])
    flags = O.gca_elements(of, k, level_range + 1, o0)
    assert 0 < flags.sum() < flags.size
    fine, kf = op_on(df, k, (D, D))
    coarse, kcv = op_on(dc, kc, (D, D))
    tng.gca_assemble_selective(fine, coarse, level_range, d0, flags)
    ref = O.gca_selective(of, oc, O.local_matrices(of, k), kc, level_range, flags)
    assert rel_err(coarse.get_local_matrices(), ref) < 1e-13
    # and through the stored-matrix apply
    x = field(oc, 3, 24)
    out = vec(dc, 3)
    coarse.apply(vec(dc, 3, x), out)
    assert rel_err(out.download(), O.epsdivdiv_stored(oc, ref, x, D, D)) < 1e-12


def test_gca_selective_rejects_mismatched_meshes(gpu_ctx):
    df, of = setup(3)
    dc, oc = setup(2)
    d0, o0 = setup(1)
    fine, kf = op_on(df, np.ones(of.shape), (D, D))
    coarse, kc = op_on(dc, np.ones(oc.shape), (D, D))
    flags = np.zeros((o0.nsd, o0.nx - 1, o0.nx - 1, o0.nr - 1), dtype=np.uint8)
    with pytest.raises(tng.TngError):
        tng.gca_assemble_selective(fine, coarse, 2, d0, flags)  # coarse is one refinement above d0
    with pytest.raises(ValueError):
        tng.gca_assemble_selective(fine, coarse, 1, d0, flags[:, :1])


_solve_cache = {}


def _solve_level4(dt, **opts):
    """Level 4 velocity / level 3 pressure, Dirichlet, layered viscosity,
    buoyancy RHS, FGMRES(30) to 1e-9."""
    key = (dt, tuple(sorted(opts.items())))
    if key not in _solve_cache:
        of = O.Mesh(4)
        r = np.broadcast_to(of.radii[:, None, None, :], of.shape)
        eta = np.where(r < 0.85, 30.0, 1.0) * np.exp(0.5 * field(of, 1, 11))
        f = field(of, 3, 12)
        S = tng.StokesSolver(2, 4, 0.55, 1.0, (D, D), tng.StokesParams.defaults(
            viscous_pc_dtype=dt, krylov_restart=30, krylov_max_iterations=400,
            krylov_relative_tolerance=1e-9, krylov_absolute_tolerance=1e-30, **opts))
        S.set_viscosity(vec(S.domain_velocity, 1, eta))
        fu = vec(S.domain_velocity, 3)
        S.buoyancy_rhs(vec(S.domain_velocity, 1, f[0]), 100.0, fu)
        u, p = vec(S.domain_velocity, 3), vec(S.domain_pressure, 1)
        rep = S.solve(u, p, fu, vec(S.domain_pressure, 1))
        _solve_cache[key] = (rep, u.download(), p.download(), of)
    return _solve_cache[key]


@pytest.mark.parametrize("dt", [tng.F64, tng.F32])
@pytest.mark.parametrize("opts", [dict(gca=1), dict(gca=2), dict(gca=1, gca_exact=1),
                                  dict(gca=1, gca_coarsest_only=1), dict(gca=2, gca_coarsest_only=1, gca_exact=1),
                                  dict(viscous_pc_smoother=1), dict(krylov_orthogonalization=1),
                                  dict(gca=1, gca_coarsest_only=1, krylov_orthogonalization=1),
                                  dict(coarse_solver=1), dict(gca=1, gca_coarsest_only=1, coarse_solver=1)])
def test_stokes_solve_options(dt, opts, gpu_ctx):
    """Galerkin preconditioners (all coarse levels, adaptive, exact product,
    coarsest level only), the damped-Jacobi smoother, the two-pass CGS2
    orthogonalisation and the direct coarsest-level solve reach the solution
    of the re-discretised FP64 default."""
    rep0, u0, p0, of = _solve_level4(tng.F64)
    rep, u, p, _ = _solve_level4(dt, **opts)
    assert rep0["converged"] == 1 and rep["converged"] == 1, (opts, rep)
    own = of.ownership_mask == 1
    assert rel_err(u[:, own], u0[:, own]) < 1e-5
    assert rel_err(p, p0) < 1e-4
