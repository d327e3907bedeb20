"""Solver layer (through the C ABI) against dense linear algebra built from the
oracle's operator applies at small levels.

- power iteration (linalg/solvers/power_iteration.hpp:13-58) vs the dense
  largest eigenvalue of D^{-1} A;
- one Chebyshev sweep (linalg/solvers/chebyshev.hpp:102-165) vs the closed-form
  Chebyshev error polynomial T_p((theta - M)/delta) / T_p(theta/delta);
- PCG (linalg/solvers/pcg.hpp:81-170) vs textbook CG on the dense matrix, and
  at level 4 vs textbook CG with the oracle's apply;
- the Stokes operator (epsilon_divdiv_stokes.hpp:55-84) vs the oracle blocks;
- V-cycle contraction, and the FGMRES Stokes solve (fgmres.hpp:109-373 with the
  block-triangular preconditioner) vs a dense direct solve, in FP64 and with
  the FP32 V-cycle.
"""
import numpy as np
import pytest

import oracle as O
import tng
from conftest import consistent, rel_err

pytestmark = pytest.mark.gpu

N, D, F = tng.NEUMANN, tng.DIRICHLET, tng.FREESLIP


def unique_index(om):
    """For every unique node id, the linear index of its OWNED copy."""
    cid = om.canonical_ids.ravel()
    own = np.nonzero(om.ownership_mask.ravel() == 1)[0]
    idx = np.empty(om.nunique, dtype=np.int64)
    idx[cid[own]] = own
    return idx


def to_unique(om, f):
    f = np.asarray(f).reshape(-1, om.nalloc)
    return f[:, unique_index(om)].ravel()


def from_unique(om, v, ncomp):
    cid = om.canonical_ids.ravel()
    v = np.asarray(v).reshape(ncomp, om.nunique)
    out = v[:, cid]
    return out.reshape(((ncomp,) if ncomp > 1 else ()) + om.shape)


def dense(om, fn, ncomp_in, om_in=None, ncomp_out=None):
    """Dense matrix on unique dofs of a linear map given on copy arrays."""
    om_in = om_in or om
    ncomp_out = ncomp_out or ncomp_in
    n_in = ncomp_in * om_in.nunique
    cols = []
    for j in range(n_in):
        e = np.zeros(n_in)
        e[j] = 1.0
        cols.append(to_unique(om, fn(from_unique(om_in, e, ncomp_in))))
    return np.array(cols).T


def field(om, ncomp, seed):
    r = np.random.default_rng(seed)
    return consistent(om, r.uniform(-1, 1, ((ncomp,) if ncomp > 1 else ()) + om.shape))


def vec(dom, ncomp, data=None, dtype=tng.F64):
    v = (tng.VectorQ1Vec if ncomp == 3 else tng.VectorQ1Scalar)(dom, dtype)
    if data is not None:
        v.upload(data)
    return v


@pytest.fixture(scope="module")
def level1():
    om = O.Mesh(1)
    dom = tng.DistributedDomain.create_uniform(1, 1, 0.55, 1.0)
    k = np.exp(field(om, 1, 5))
    A = dense(om, lambda x: om.epsdivdiv(k, x, D, D, path="fast"), 3)
    diag = to_unique(om, om.epsdivdiv(k, np.ones((3,) + om.shape), D, D, diagonal=True, path="fast"))
    return om, dom, k, A, diag


def gpu_op(dom, k, bcs=(D, D)):
    kv = vec(dom, 1, k)
    return tng.EpsilonDivDivKerngen(dom, kv, bcs), kv


def inv_diag_vec(dom, op):
    ones, d = vec(dom, 3), vec(dom, 3)
    ones.fill(1.0)
    op.set_diagonal(True)
    op.apply(ones, d)
    op.set_diagonal(False)
    tng.invert_entries(d)
    return d


def test_power_iteration(level1, gpu_ctx):
    om, dom, k, A, diag = level1
    M = A / diag[:, None]
    lmax = np.max(np.abs(np.linalg.eigvals(M)))
    op, kv = gpu_op(dom, k)
    invd = inv_diag_vec(dom, op)
    assert rel_err(to_unique(om, invd.download()), 1.0 / diag) < 1e-13
    est10 = tng.power_iteration(op, invd, 10)
    est400 = tng.power_iteration(op, invd, 400)
    assert 0.3 * lmax < est10 < 1.2 * lmax
    assert abs(est400 - lmax) < 2e-3 * lmax
    # the seeded start makes the estimate reproducible
    assert tng.power_iteration(op, invd, 10) == est10


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_chebyshev_closed_form(order, level1, gpu_ctx):
    """One sweep from x = 0 on A x = A x*: the error is the Chebyshev
    polynomial T_p((theta I - M)/delta) / T_p(theta/delta) applied to x*."""
    om, dom, k, A, diag = level1
    M = A / diag[:, None]
    est = 0.9 * np.max(np.abs(np.linalg.eigvals(M)))
    lmax, lmin = 1.5 * est, 0.1 * est
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    xs = field(om, 3, 7)
    xu = to_unique(om, xs)
    X = lambda v: (theta * v - M @ v) / delta  # noqa: E731
    t_prev, t = xu, X(xu)
    s_prev, s = 1.0, theta / delta
    for _ in range(order - 1):
        t_prev, t = t, 2 * X(t) - t_prev
        s_prev, s = s, 2 * (theta / delta) * s - s_prev
    err_exp = (t if order > 1 else X(xu)) / (s if order > 1 else theta / delta)
    op, kv = gpu_op(dom, k)
    invd = inv_diag_vec(dom, op)
    b = vec(dom, 3)
    op.apply(vec(dom, 3, xs), b)
    x = vec(dom, 3)
    tng.chebyshev_smooth(op, invd, x, b, order=order, iterations=1, lambda_max_estimate=est)
    err = xu - to_unique(om, x.download())
    assert rel_err(err, err_exp) < 1e-11


@pytest.mark.parametrize("iterations,omega", [(1, 1.0), (3, 0.7)])
def test_jacobi_closed_form(iterations, omega, level1, gpu_ctx):
    """Damped Jacobi (jacobi.hpp:16-60) from a nonzero start on A x = A x*:
    the error after n sweeps is (I - omega D^{-1} A)^n (x0 - x*)."""
    om, dom, k, A, diag = level1
    M = A / diag[:, None]
    xs, x0 = field(om, 3, 9), field(om, 3, 10)
    e = to_unique(om, x0) - to_unique(om, xs)
    for _ in range(iterations):
        e = e - omega * (M @ e)
    op, kv = gpu_op(dom, k)
    invd = inv_diag_vec(dom, op)
    b = vec(dom, 3)
    op.apply(vec(dom, 3, xs), b)
    x = vec(dom, 3, x0)
    tng.jacobi_smooth(op, invd, x, b, iterations=iterations, omega=omega)
    err = to_unique(om, x.download()) - to_unique(om, xs)
    assert rel_err(err, e) < 1e-11


def test_pcg_matches_textbook_cg(level1, gpu_ctx):
    om, dom, k, A, diag = level1
    xs = field(om, 3, 8)
    bu = A @ to_unique(om, xs)
    op, kv = gpu_op(dom, k)
    b = vec(dom, 3)
    op.apply(vec(dom, 3, xs), b)
    for kmax in (5, 15):
        x = vec(dom, 3)
        it, rel = tng.pcg_solve(op, x, b, max_iterations=kmax, rtol=0.0, atol=0.0)
        # textbook CG, x0 = 0
        xr = np.zeros_like(bu)
        r = bu.copy()
        p = r.copy()
        for _ in range(kmax):
            ap = A @ p
            a = (r @ r) / (p @ ap)
            xr += a * p
            rn = r - a * ap
            p = rn + (rn @ rn) / (r @ r) * p
            r = rn
        assert it == kmax
        assert rel_err(to_unique(om, x.download()), xr) < 1e-9
        assert abs(rel - np.linalg.norm(r) / np.linalg.norm(bu)) < 1e-8
    x = vec(dom, 3)
    it, rel = tng.pcg_solve(op, x, b, max_iterations=500, rtol=1e-12, atol=0.0)
    assert rel <= 1e-12
    assert rel_err(to_unique(om, x.download()), to_unique(om, xs)) < 1e-9


def test_cg_level4_dirichlet(gpu_ctx):
    """The CPU-sized configuration's solve: level-4 shell, eta = exp(z) with z
    ~ N(0, 1) per node, Dirichlet on both surfaces, CG from x0 = 0. The first
    20 iterates equal textbook CG in numpy on unique dofs with the oracle's
    apply as the matvec; run to 1e-10, the true residual b - A x of the result
    is 1e-10 too."""
    om = O.Mesh(4)
    dom = tng.DistributedDomain.create_uniform(4, 4, 0.55, 1.0)
    k = np.exp(consistent(om, np.random.default_rng(60).normal(size=om.shape)))
    xs = field(om, 3, 61)
    op, kv = gpu_op(dom, k)
    b = vec(dom, 3)
    op.apply(vec(dom, 3, xs), b)
    bu = to_unique(om, b.download())

    def A(pu):
        return to_unique(om, om.epsdivdiv(k, from_unique(om, pu, 3), D, D, path="fast"))

    kmax = 20
    xr, r = np.zeros_like(bu), bu.copy()
    p = r.copy()
    for _ in range(kmax):
        ap = A(p)
        a = (r @ r) / (p @ ap)
        xr += a * p
        rn = r - a * ap
        p = rn + (rn @ rn) / (r @ r) * p
        r = rn
    x = vec(dom, 3)
    it, rel = tng.pcg_solve(op, x, b, max_iterations=kmax, rtol=0.0, atol=0.0)
    assert it == kmax
    assert rel_err(to_unique(om, x.download()), xr) < 1e-9
    assert abs(rel - np.linalg.norm(r) / np.linalg.norm(bu)) < 1e-8

    x = vec(dom, 3)
    it, rel = tng.pcg_solve(op, x, b, max_iterations=5000, rtol=1e-10, atol=0.0)
    assert rel <= 1e-10, (it, rel)
    ax = vec(dom, 3)
    op.apply(x, ax)
    true_rel = np.linalg.norm(bu - to_unique(om, ax.download())) / np.linalg.norm(bu)
    assert true_rel < 1e-9, (it, true_rel)


# ------------------------------------------------------------------ Stokes
def stokes_solver(minl, maxl, bcs, **kw):
    return tng.StokesSolver(minl, maxl, 0.55, 1.0, bcs, tng.StokesParams.defaults(**kw))


@pytest.mark.parametrize("bcs", [(D, D), (F, F), (F, D)])
def test_stokes_apply(bcs, gpu_ctx):
    S = stokes_solver(1, 3, bcs)
    of, oc = O.Mesh(3), O.Mesh(2)
    u, p = field(of, 3, 9), field(oc, 1, 10)
    ou, op_ = vec(S.domain_velocity, 3), vec(S.domain_pressure, 1)
    S.apply(vec(S.domain_velocity, 3, u), vec(S.domain_pressure, 1, p), ou, op_)
    k = np.ones(of.shape)
    ru = of.epsdivdiv(k, u, *bcs, apply_mode=O.REPLACE, comm=O.SKIP, path="fast")
    ru = O.gradient(of, oc, p, *bcs, apply_mode=O.ADD, comm=O.COMMUNICATE, dst=ru)
    rp = O.divergence(of, oc, u, *bcs)
    assert rel_err(ou.download(), ru) < 1e-12
    assert rel_err(op_.download(), rp) < 1e-12


@pytest.mark.parametrize("dt,tol", [(tng.F64, 1e-10), (tng.F32, 2e-5)])
def test_vcycle_matches_restatement(dt, tol, gpu_ctx):
    """One V-cycle (velocity levels 1..3, Dirichlet, log-normal viscosity
    injected to the coarse levels, default smoother and eigenvalue estimates)
    against the numpy restatement on the oracle's operators."""
    from mg_reference import VCycle

    of = O.Mesh(3)
    eta = np.exp(0.5 * field(of, 1, 21))
    S = stokes_solver(1, 3, (D, D), viscous_pc_dtype=dt)
    S.set_viscosity(vec(S.domain_velocity, 1, eta))
    b = field(of, 3, 22)
    x = vec(S.domain_velocity, 3)
    S.precondition_velocity(x, vec(S.domain_velocity, 3, b))
    ref = VCycle(1, 3, eta, (D, D))(np.zeros_like(b), b)
    assert rel_err(x.download(), ref) < tol


@pytest.mark.parametrize("dt", [tng.F64, tng.F32])
@pytest.mark.parametrize("bcs", [(D, D), (F, F), (D, F)])
def test_fused_chebyshev_matches_separate_kernels(dt, bcs, gpu_ctx):
    """Chebyshev sub-steps fused into the one-column apply kernels (levels with
    64 and 128 radial cells) and the shared-node update after the halo sum give
    bit-identical V-cycles and FGMRES iterates to the separate apply, halo sum
    and update kernels."""
    ov, op_ = O.Mesh(5, 7), O.Mesh(4, 6)
    eta = np.exp(0.5 * field(ov, 1, 31))
    b = field(ov, 3, 32)
    fp = field(op_, 1, 33)
    out = {}
    for fuse in ("0", "1"):
        with tng.default_context().tuning(cheb_fuse=int(fuse)):
            S = tng.StokesSolver(3, 5, 0.55, 1.0, bcs, tng.StokesParams.defaults(
                viscous_pc_dtype=dt, krylov_max_iterations=6, krylov_restart=4, krylov_relative_tolerance=1e-30),
                radial_extra_levels=2)
        S.set_viscosity(vec(S.domain_velocity, 1, eta))
        x = vec(S.domain_velocity, 3)
        bv = vec(S.domain_velocity, 3, b)
        S.precondition_velocity(x, bv)
        S.precondition_velocity(x, bv)  # from a nonzero start
        u, p = vec(S.domain_velocity, 3), vec(S.domain_pressure, 1)
        S.solve(u, p, bv, vec(S.domain_pressure, 1, fp))
        out[fuse] = (x.download(), u.download(), p.download())
    for name, a, c in zip(("V-cycle", "FGMRES u", "FGMRES p"), out["0"], out["1"]):
        bad = np.argwhere(a != c)
        assert bad.size == 0, (name, len(bad), bad[:8].tolist(), float(np.abs(a - c).max() / np.abs(a).max()))


def test_fgmres_zero_rhs_returns_immediately(gpu_ctx):
    """fgmres.hpp:130-133: a zero right-hand side converges with zero iterations and u = 0."""
    S = stokes_solver(1, 3, (D, D))
    u, p = vec(S.domain_velocity, 3), vec(S.domain_pressure, 1)
    u.fill(0.0)
    rep = S.solve(u, p, vec(S.domain_velocity, 3), vec(S.domain_pressure, 1))
    assert rep["converged"] == 1 and rep["iterations"] == 0
    assert not u.download().any() and not p.download().any()


def test_fgmres_skips_nonfinite_preconditioner(gpu_ctx):
    """fgmres.hpp:189-202: a preconditioner output with NaN/Inf is replaced by the
    Arnoldi vector itself.  Forced from the host: one interior viscosity value
    above the FP32 range leaves the FP64 operator finite but overflows the FP32
    V-cycle, so every step takes the guard and FGMRES still reduces the
    residual (unpreconditioned)."""
    of = O.Mesh(3)
    eta = np.ones(of.shape)
    eta[3, 4, 4, 4] = 1e39  # interior node of one subdomain: a single copy
    S = stokes_solver(1, 3, (D, D), viscous_pc_dtype=tng.F32, krylov_max_iterations=6, krylov_restart=10)
    S.set_viscosity(vec(S.domain_velocity, 1, eta))
    u, p = vec(S.domain_velocity, 3), vec(S.domain_pressure, 1)
    rep = S.solve(u, p, vec(S.domain_velocity, 3, field(of, 3, 41)), vec(S.domain_pressure, 1))
    assert rep["nonfinite_preconditioner"] == rep["iterations"] == 6, rep
    assert rep["nonfinite_residual"] == 0
    assert np.isfinite(rep["final_residual"]) and rep["final_residual"] < rep["initial_residual"]
    assert np.isfinite(u.download()).all()


def test_vcycle_reduces_error(gpu_ctx):
    """V-cycles with the reference's default smoother reduce the error; the
    contraction is limited by the low-energy modes of the one-point wedge
    quadrature that neither the smoother nor the coarse space resolves
    (the numpy restatement shows the same rates)."""
    of = O.Mesh(4)
    S = stokes_solver(2, 4, (D, D))
    dom = S.domain_velocity
    xs = field(of, 3, 11)
    kv = vec(dom, 1)
    kv.fill(1.0)
    A = tng.EpsilonDivDivKerngen(dom, kv, (D, D))
    b = vec(dom, 3)
    A.apply(vec(dom, 3, xs), b)
    x = vec(dom, 3)
    errs = []
    for _ in range(4):
        S.precondition_velocity(x, b)
        errs.append(np.linalg.norm(x.download() - xs))
    assert errs[0] < 0.2 * np.linalg.norm(xs)
    assert all(errs[i + 1] < 0.9 * errs[i] for i in range(3)), errs


@pytest.fixture(scope="module")
def stokes_dense():
    """Dense K on unique dofs: velocity level 2, pressure level 1, Dirichlet."""
    of, oc = O.Mesh(2), O.Mesh(1)
    eta = np.exp(0.5 * field(of, 1, 12))
    nu, npr = 3 * of.nunique, oc.nunique

    def K_u(u):
        out = of.epsdivdiv(eta, u, D, D, comm=O.SKIP, path="fast")
        return of.halo_sum(out), O.divergence(of, oc, u, D, D)

    Kuu = dense(of, lambda u: K_u(u)[0], 3)
    Kpu = dense(oc, lambda u: K_u(u)[1], 3, om_in=of, ncomp_out=1)
    Kup = dense(of, lambda p: O.gradient(of, oc, p, D, D), 1, om_in=oc, ncomp_out=3)
    K = np.block([[Kuu, Kup], [Kpu, np.zeros((npr, npr))]])
    return of, oc, eta, K, nu, npr


@pytest.mark.parametrize("orth", [0, 1])
@pytest.mark.parametrize("dt", [tng.F64, tng.F32])
def test_stokes_solve_vs_dense(dt, orth, stokes_dense, gpu_ctx):
    of, oc, eta, K, nu, npr = stokes_dense
    S = stokes_solver(1, 2, (D, D), krylov_restart=30, krylov_max_iterations=600, krylov_relative_tolerance=1e-10,
                      krylov_absolute_tolerance=1e-30, viscous_pc_dtype=dt, krylov_orthogonalization=orth)
    S.set_viscosity(vec(S.domain_velocity, 1, eta))
    xs_u, xs_p = field(of, 3, 13), field(oc, 1, 14)
    fu, fp = vec(S.domain_velocity, 3), vec(S.domain_pressure, 1)
    S.apply(vec(S.domain_velocity, 3, xs_u), vec(S.domain_pressure, 1, xs_p), fu, fp)
    # dense reference: minimum-norm solution, then the pressure mean removed
    f = np.concatenate([to_unique(of, fu.download()), to_unique(oc, fp.download())])
    sol = np.linalg.lstsq(K, f, rcond=None)[0]
    u_ref, p_ref = sol[:nu], sol[nu:] - sol[nu:].mean()
    u, p = vec(S.domain_velocity, 3), vec(S.domain_pressure, 1)
    rep = S.solve(u, p, fu, fp)
    assert rep["converged"] == 1, rep
    assert rep["relative_residual"] <= 1e-10
    assert rel_err(to_unique(of, u.download()), u_ref) < 1e-8
    assert rel_err(to_unique(oc, p.download()), p_ref) < 1e-6
    # the velocity of the manufactured solution is recovered (pressure up to a constant)
    assert rel_err(to_unique(of, u.download()), to_unique(of, xs_u)) < 1e-8


@pytest.mark.parametrize("bcs,rtol", [((D, D), 1e-8), ((D, F), 1e-6)])
def test_stokes_solve_fp32_vcycle_matches_fp64(bcs, rtol, gpu_ctx):
    """Velocity level 4 / pressure level 3, buoyancy RHS, viscosity jump: the
    FP32 V-cycle preconditioner reaches the same FP64 residual and the same
    solution as the FP64 one.  With a free-slip surface the constant pressure
    is only nearly in the kernel of the gradient (the free-slip rows project
    with the nodal radial normal, not the discrete face normal), which leaves
    a near-singular mode that slows the restarted FGMRES below about 1e-7."""
    of, oc = O.Mesh(4), O.Mesh(3)
    r = np.broadcast_to(of.radii[:, None, None, :], of.shape)
    eta = np.where(r < 0.85, 10.0, 1.0) * np.exp(0.5 * field(of, 1, 15))
    T = field(of, 1, 16)
    sols = {}
    for dt in (tng.F64, tng.F32):
        S = stokes_solver(2, 4, bcs, krylov_restart=30, krylov_max_iterations=400, krylov_relative_tolerance=rtol,
                          krylov_absolute_tolerance=1e-30, viscous_pc_dtype=dt, viscous_pc_chebyshev_order=3)
        S.set_viscosity(vec(S.domain_velocity, 1, eta))
        fu = vec(S.domain_velocity, 3)
        S.buoyancy_rhs(vec(S.domain_velocity, 1, T), 100.0, fu)
        fp = vec(S.domain_pressure, 1)
        u, p = vec(S.domain_velocity, 3), vec(S.domain_pressure, 1)
        rep = S.solve(u, p, fu, fp)
        assert rep["converged"] == 1 and rep["relative_residual"] <= rtol, rep
        sols[dt] = (u.download(), p.download(), rep)
    (u64, p64, r64), (u32, p32, r32) = sols[tng.F64], sols[tng.F32]
    own = of.ownership_mask == 1
    assert rel_err(u32[:, own], u64[:, own]) < 1e3 * rtol
    assert r32["iterations"] <= 2 * r64["iterations"] + 5


def test_stokes_solve_freeslip_buoyancy(gpu_ctx):
    """Free slip on both surfaces with the rotation penalty: the buoyancy RHS
    (normal forcing, no rotational component up to discretisation) converges
    until the rotations' eps-scale eigenvalues dominate (about 1e-6)."""
    of = O.Mesh(4)
    S = stokes_solver(2, 4, (F, F), krylov_restart=30, krylov_max_iterations=200, krylov_relative_tolerance=1e-6,
                      krylov_absolute_tolerance=1e-30)
    fu = vec(S.domain_velocity, 3)
    S.buoyancy_rhs(vec(S.domain_velocity, 1, field(of, 1, 17)), 100.0, fu)
    u, p = vec(S.domain_velocity, 3), vec(S.domain_pressure, 1)
    rep = S.solve(u, p, fu, vec(S.domain_pressure, 1))
    assert rep["converged"] == 1, rep


def test_rotation_correction_reaches_the_same_solution(gpu_ctx):
    """Free slip on both surfaces: with the rigid-rotation correction after the
    block preconditioner (every preconditioned step made exact on the
    rotations, Z^T (v_j - K z_j) = 0) the solve converges to the same solution
    as without, up to rigid rotations: the penalty makes them near-null modes,
    which a 1e-6 residual leaves loose."""
    of = O.Mesh(4)
    r = np.broadcast_to(of.radii[:, None, None, :], of.shape)
    eta = np.where(r < 0.85, 10.0, 1.0) * np.exp(0.5 * field(of, 1, 43))
    T = field(of, 1, 44)
    out = {}
    for rc in (0, 1):
        S = stokes_solver(2, 4, (F, F), krylov_restart=30, krylov_max_iterations=300, krylov_relative_tolerance=1e-6,
                          krylov_absolute_tolerance=1e-30, rotation_correction=rc)
        S.set_viscosity(vec(S.domain_velocity, 1, eta))
        fu = vec(S.domain_velocity, 3)
        S.buoyancy_rhs(vec(S.domain_velocity, 1, T), 100.0, fu)
        u, p = vec(S.domain_velocity, 3), vec(S.domain_pressure, 1)
        rep = S.solve(u, p, fu, vec(S.domain_pressure, 1))
        assert rep["converged"] == 1, (rc, rep)
        out[rc] = (u.download(),)
    own = of.ownership_mask == 1
    X = of.coords[:, :, :, None, :] * of.radii[:, None, None, :, None]
    Z = np.stack([np.cross(np.eye(3)[i], X[own]).ravel() for i in range(3)], axis=1)
    Q, _ = np.linalg.qr(Z)
    u0, u1 = (np.moveaxis(out[rc][0], 0, -1)[own].ravel() for rc in (0, 1))
    d = (u1 - u0) - Q @ (Q.T @ (u1 - u0))
    assert np.linalg.norm(d) < 1e-4 * np.linalg.norm(u0), np.linalg.norm(d) / np.linalg.norm(u0)


@pytest.mark.parametrize("orth", [0, 1])
def test_rotation_correction_with_the_fused_smoother(orth, gpu_ctx):
    """64 radial cells on the top level, where the fused Chebyshev sub-steps
    exchange the storage of the vectors they update: the rotation correction
    (and, with CGS2, the batched solution update) must follow the
    preconditioned directions to their current storage.  The fused and the
    separate smoother kernels give the same bits, so the two solves must
    agree exactly."""
    ov = O.Mesh(4, 6)
    r = np.broadcast_to(ov.radii[:, None, None, :], ov.shape)
    eta = np.where(r < 0.85, 10.0, 1.0) * np.exp(0.5 * field(ov, 1, 47))
    T = field(ov, 1, 48)
    out = {}
    for fuse in (0, 1):
        with tng.default_context().tuning(cheb_fuse=fuse):
            S = tng.StokesSolver(2, 4, 0.55, 1.0, (F, F), tng.StokesParams.defaults(
                krylov_restart=8, krylov_max_iterations=20, krylov_relative_tolerance=1e-30,
                krylov_absolute_tolerance=1e-30, rotation_correction=1, krylov_orthogonalization=orth),
                radial_extra_levels=2)
        S.set_viscosity(vec(S.domain_velocity, 1, eta))
        fu = vec(S.domain_velocity, 3)
        S.buoyancy_rhs(vec(S.domain_velocity, 1, T), 100.0, fu)
        u, p = vec(S.domain_velocity, 3), vec(S.domain_pressure, 1)
        rep = S.solve(u, p, fu, vec(S.domain_pressure, 1), history_capacity=32)
        out[fuse] = (u.download(), p.download(), np.array(rep["history"]))
    for name, a, c in zip(("u", "p", "history"), out[0], out[1]):
        bad = np.argwhere(a != c)
        assert bad.size == 0, (name, len(bad), float(np.abs(a - c).max() / np.abs(a).max()))


@pytest.mark.parametrize("dt,bcs", [(tng.F64, (F, F)), (tng.F32, (D, F))])
def test_direct_coarse_solve_free_slip(dt, bcs, gpu_ctx):
    """The direct coarsest-level solve (dense inverse of the level-2 operator
    with its free-slip rows and, free slip on both surfaces, the rotation
    penalty) reaches the PCG coarse solve's solution up to rigid rotations,
    in no more FGMRES iterations.  (With free slip on both surfaces and the
    FP32 V-cycle the PCG variant stalls above 1e-5 here.)"""
    of = O.Mesh(4)
    r = np.broadcast_to(of.radii[:, None, None, :], of.shape)
    eta = np.where(r < 0.85, 10.0, 1.0) * np.exp(0.5 * field(of, 1, 45))
    T = field(of, 1, 46)
    out = {}
    for cs in (0, 1):
        S = stokes_solver(2, 4, bcs, krylov_restart=30, krylov_max_iterations=300, krylov_relative_tolerance=1e-6,
                          krylov_absolute_tolerance=1e-30, viscous_pc_dtype=dt, coarse_solver=cs)
        S.set_viscosity(vec(S.domain_velocity, 1, eta))
        fu = vec(S.domain_velocity, 3)
        S.buoyancy_rhs(vec(S.domain_velocity, 1, T), 100.0, fu)
        u, p = vec(S.domain_velocity, 3), vec(S.domain_pressure, 1)
        rep = S.solve(u, p, fu, vec(S.domain_pressure, 1))
        assert rep["converged"] == 1, (cs, rep)
        out[cs] = (u.download(), rep["iterations"])
    assert out[1][1] <= out[0][1] + 2, (out[0][1], out[1][1])
    own = of.ownership_mask == 1
    X = of.coords[:, :, :, None, :] * of.radii[:, None, None, :, None]
    Z = np.stack([np.cross(np.eye(3)[i], X[own]).ravel() for i in range(3)], axis=1)
    Q, _ = np.linalg.qr(Z)
    u0, u1 = (np.moveaxis(out[cs][0], 0, -1)[own].ravel() for cs in (0, 1))
    d = (u1 - u0) - Q @ (Q.T @ (u1 - u0))
    assert np.linalg.norm(d) < 1e-4 * np.linalg.norm(u0), np.linalg.norm(d) / np.linalg.norm(u0)


def test_enforce_rhs_first_boundary_only(gpu_ctx):
    """Dirichlet on both surfaces: only the CMB rows are zeroed, as the
    reference's get_shell_boundary_flag returns the first match."""
    S = stokes_solver(1, 2, (D, D))
    of = O.Mesh(2)
    f = field(of, 3, 17)
    fu = vec(S.domain_velocity, 3, f)
    S.enforce_rhs(fu)
    g = fu.download()
    assert np.all(g[..., 0] == 0)
    assert np.array_equal(g[..., 1:], f[..., 1:])
    S2 = stokes_solver(1, 2, (D, F))
    fu = vec(S2.domain_velocity, 3, f)
    S2.enforce_rhs(fu)
    g = fu.download()
    n = of.coords / np.linalg.norm(of.coords, axis=-1, keepdims=True)
    assert np.all(g[..., 0] == 0)
    assert np.abs(np.einsum("csxy,sxyc->sxy", g[..., -1], n)).max() < 1e-15
    assert np.array_equal(g[..., 1:-1], f[..., 1:-1])


@pytest.mark.parametrize("bcs,rtol,restart", [((D, D), 1e-9, 30), ((D, D), 1e-9, 8), ((F, D), 1e-7, 30)])
def test_cgs2_matches_mgs(bcs, rtol, restart, gpu_ctx):
    """Two-pass CGS2 orthogonalisation (one reduction per Arnoldi step, the
    re-orthogonalisation delayed by one step) against the reference's modified
    Gram-Schmidt: same iteration count within one, the same residual history
    to 1e-6 relative while the residual is above 1e3 rtol, the same solution."""
    of = O.Mesh(4)
    r = np.broadcast_to(of.radii[:, None, None, :], of.shape)
    eta = np.where(r < 0.85, 10.0, 1.0) * np.exp(0.5 * field(of, 1, 41))
    T = field(of, 1, 42)
    out = {}
    for orth in (0, 1):
        S = stokes_solver(2, 4, bcs, krylov_restart=restart, krylov_max_iterations=600, krylov_relative_tolerance=rtol,
                          krylov_absolute_tolerance=1e-30, viscous_pc_chebyshev_order=3, krylov_orthogonalization=orth)
        S.set_viscosity(vec(S.domain_velocity, 1, eta))
        fu = vec(S.domain_velocity, 3)
        S.buoyancy_rhs(vec(S.domain_velocity, 1, T), 100.0, fu)
        u, p = vec(S.domain_velocity, 3), vec(S.domain_pressure, 1)
        rep = S.solve(u, p, fu, vec(S.domain_pressure, 1), history_capacity=2000)
        assert rep["converged"] == 1 and rep["relative_residual"] <= rtol, (orth, rep)
        out[orth] = (rep, u.download())
    (r0, u0), (r1, u1) = out[0], out[1]
    assert abs(r0["iterations"] - r1["iterations"]) <= 1, (r0["iterations"], r1["iterations"])
    h0, h1 = np.array(r0["history"]), np.array(r1["history"])
    n = min(len(h0), len(h1))
    keep = h0[:n] > 1e3 * rtol * h0[0]
    assert np.all(np.abs(h1[:n][keep] - h0[:n][keep]) <= 1e-6 * h0[:n][keep])
    own = of.ownership_mask == 1
    assert rel_err(u1[:, own], u0[:, own]) < 1e2 * rtol
