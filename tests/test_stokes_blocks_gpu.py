"""Parity of the HIP divergence B, gradient B^T, prolongation P, restriction R,
lumped KMass and the vector kernels (through the C ABI) with the CPU oracle.

Reference: divergence_kerngen.hpp:206-257, gradient_kerngen.hpp:207-258,
prolongation_constant.hpp:102-191, restriction_constant.hpp:133-250,
kmass.hpp:72-181, linalg/vector.hpp:72-197, linalg/vector_q1.hpp:62-126.
FP64 bar: 1e-12 relative (||diff||_2 / ||ref||_2); FP32 bar: 2e-6."""
import numpy as np
import pytest

import oracle as O
import tng
from conftest import consistent, rel_err

pytestmark = pytest.mark.gpu

N, D, F = tng.NEUMANN, tng.DIRICHLET, tng.FREESLIP
_cache = {}


def levels(L, ls=0, rs=0, R=None):
    R = L if R is None else R
    key = (L, ls, rs, R)
    if key not in _cache:
        df = tng.DistributedDomain.create_uniform(L, R, 0.55, 1.0, ls, rs)
        dc = tng.DistributedDomain.create_uniform(L - 1, R - 1, 0.55, 1.0, ls, rs)
        of = O.Mesh(L, R, 0.55, 1.0, ls, rs)
        oc = O.Mesh(L - 1, R - 1, 0.55, 1.0, ls, rs)
        _cache[key] = (df, dc, of, oc)
    return _cache[key]


def field(om, ncomp, seed):
    r = np.random.default_rng(seed)
    return consistent(om, r.uniform(-1, 1, ((ncomp,) if ncomp > 1 else ()) + om.shape))


def vec(dom, ncomp, data=None, dtype=tng.F64):
    v = (tng.VectorQ1Vec if ncomp == 3 else tng.VectorQ1Scalar)(dom, dtype)
    if data is not None:
        v.upload(data)
    return v


BC_SETS = [(N, N), (D, D), (F, F), (D, F), (F, N)]
SHAPES = [(3, 0, 0), (4, 0, 0), (3, 1, 1), (4, 1, 0)]
# tall columns (65, 129 and 257 nodes per subdomain column): the LDS-staged
# kernels and the one-wave-per-column gradient (1, 2 and 4 cells per lane)
TALL = [(3, 0, 0, 6), (3, 0, 0, 7), (3, 1, 1, 7), (2, 0, 0, 7), (2, 0, 0, 8)]


@pytest.mark.parametrize("bcs", BC_SETS)
@pytest.mark.parametrize("L,ls,rs,R", [s + (None,) for s in SHAPES] + TALL)
def test_divergence(bcs, L, ls, rs, R, gpu_ctx):
    df, dc, of, oc = levels(L, ls, rs, R)
    u = field(of, 3, 1)
    op = tng.DivergenceKerngen(df, dc, bcs)
    out = vec(dc, 1)
    op.apply(vec(df, 3, u), out)
    ref = O.divergence(of, oc, u, *bcs)
    assert rel_err(out.download(), ref) < 1e-12


@pytest.mark.parametrize("bcs", BC_SETS)
@pytest.mark.parametrize("L,ls,rs,R", [s + (None,) for s in SHAPES] + TALL)
def test_gradient(bcs, L, ls, rs, R, gpu_ctx):
    df, dc, of, oc = levels(L, ls, rs, R)
    p = field(oc, 1, 2)
    op = tng.GradientKerngen(df, dc, bcs)
    out = vec(df, 3)
    op.apply(vec(dc, 1, p), out)
    ref = O.gradient(of, oc, p, *bcs)
    assert rel_err(out.download(), ref) < 1e-12


@pytest.mark.parametrize("shape", [(3, 1, 0, None), (3, 1, 0, 7), (2, 0, 1, 7)])
@pytest.mark.parametrize("mode,comm", [(tng.ADD, tng.COMMUNICATE_ADDITIVELY), (tng.ADD, tng.SKIP_COMMUNICATION),
                                       (tng.REPLACE, tng.SKIP_COMMUNICATION)])
def test_divergence_gradient_modes(mode, comm, shape, gpu_ctx):
    df, dc, of, oc = levels(*shape)
    u, p = field(of, 3, 3), field(oc, 1, 4)
    b0, g0 = field(oc, 1, 5), field(of, 3, 6)
    B = tng.DivergenceKerngen(df, dc, (D, F))
    B.set_operator_apply_and_communication_modes(mode, comm)
    Bt = tng.GradientKerngen(df, dc, (F, D))
    Bt.set_operator_apply_and_communication_modes(mode, comm)
    ob, og = vec(dc, 1, b0), vec(df, 3, g0)
    B.apply(vec(df, 3, u), ob)
    Bt.apply(vec(dc, 1, p), og)
    rb = O.divergence(of, oc, u, D, F, apply_mode=mode, comm=comm, dst=b0)
    rg = O.gradient(of, oc, p, F, D, apply_mode=mode, comm=comm, dst=g0)
    assert rel_err(ob.download(), rb) < 1e-12
    assert rel_err(og.download(), rg) < 1e-12


def test_divergence_gradient_level6_shape(gpu_ctx):
    """Velocity L5 -> pressure L4 (the velocity/pressure pair below the
    L7/L6 solve), free slip on both surfaces."""
    df, dc, of, oc = levels(5)
    u, p = field(of, 3, 7), field(oc, 1, 8)
    ob, og = vec(dc, 1), vec(df, 3)
    tng.DivergenceKerngen(df, dc, (F, F)).apply(vec(df, 3, u), ob)
    tng.GradientKerngen(df, dc, (F, F)).apply(vec(dc, 1, p), og)
    assert rel_err(ob.download(), O.divergence(of, oc, u, F, F)) < 1e-12
    assert rel_err(og.download(), O.gradient(of, oc, p, F, F)) < 1e-12


def test_divergence_gradient_fp32(gpu_ctx):
    df, dc, of, oc = levels(4)
    u, p = field(of, 3, 9), field(oc, 1, 10)
    ob, og = vec(dc, 1, dtype=tng.F32), vec(df, 3, dtype=tng.F32)
    tng.DivergenceKerngen(df, dc, (D, F)).apply(vec(df, 3, u, tng.F32), ob)
    tng.GradientKerngen(df, dc, (D, F)).apply(vec(dc, 1, p, tng.F32), og)
    assert rel_err(ob.download(), O.divergence(of, oc, u, D, F)) < 2e-6
    assert rel_err(og.download(), O.gradient(of, oc, p, D, F)) < 2e-6


@pytest.mark.parametrize("L,ls,rs,R", [(3, 0, 0, 6), (3, 1, 1, 7)])
def test_divergence_gradient_fp32_tall(L, ls, rs, R, gpu_ctx):
    df, dc, of, oc = levels(L, ls, rs, R)
    u, p = field(of, 3, 13), field(oc, 1, 14)
    ob, og = vec(dc, 1, dtype=tng.F32), vec(df, 3, dtype=tng.F32)
    tng.DivergenceKerngen(df, dc, (F, D)).apply(vec(df, 3, u, tng.F32), ob)
    tng.GradientKerngen(df, dc, (F, D)).apply(vec(dc, 1, p, tng.F32), og)
    assert rel_err(ob.download(), O.divergence(of, oc, u, F, D)) < 2e-6
    assert rel_err(og.download(), O.gradient(of, oc, p, F, D)) < 2e-6


@pytest.mark.parametrize("ncomp", [1, 3])
@pytest.mark.parametrize("L,ls,rs", SHAPES)
def test_prolongation_restriction(ncomp, L, ls, rs, gpu_ctx):
    df, dc, of, oc = levels(L, ls, rs)
    xc, xf = field(oc, ncomp, 11), field(of, ncomp, 12)
    P = tng.ProlongationConstant(dc, df)
    R = tng.RestrictionConstant(df, dc)
    pf, rc = vec(df, ncomp), vec(dc, ncomp)
    P.apply(vec(dc, ncomp, xc), pf)
    R.apply(vec(df, ncomp, xf), rc)
    assert rel_err(pf.download(), O.prolongate(oc, of, xc)) < 1e-15
    assert rel_err(rc.download(), O.restrict(of, oc, xf)) < 1e-15
    # Add mode (the V-cycle's coarse-grid correction)
    base = field(of, ncomp, 13)
    pa = vec(df, ncomp, base)
    P.set_operator_apply_and_communication_modes(tng.ADD, tng.SKIP_COMMUNICATION)
    P.apply(vec(dc, ncomp, xc), pa)
    assert rel_err(pa.download(), O.prolongate(oc, of, xc, apply_mode=O.ADD, dst=base)) < 1e-15


def test_prolongation_restriction_fp32(gpu_ctx):
    df, dc, of, oc = levels(4)
    xc, xf = field(oc, 3, 14), field(of, 3, 15)
    pf, rc = vec(df, 3, dtype=tng.F32), vec(dc, 3, dtype=tng.F32)
    tng.ProlongationConstant(dc, df).apply(vec(dc, 3, xc, tng.F32), pf)
    tng.RestrictionConstant(df, dc).apply(vec(df, 3, xf, tng.F32), rc)
    assert rel_err(pf.download(), O.prolongate(oc, of, xc)) < 2e-7
    assert rel_err(rc.download(), O.restrict(of, oc, xf)) < 1e-6


@pytest.mark.parametrize("L,ls,rs", [(3, 0, 0), (4, 1, 1), (5, 0, 0)])
def test_kmass_lumped(L, ls, rs, gpu_ctx):
    df, dc, of, oc = levels(L, ls, rs)
    k = np.exp(field(oc, 1, 16))
    x = field(oc, 1, 17)
    kv = vec(dc, 1, k)
    M = tng.KMassLumped(dc, kv)
    out = vec(dc, 1)
    M.apply(vec(dc, 1, x), out)
    assert rel_err(out.download(), O.kmass_lumped(oc, k, x)) < 1e-12


def test_kmass_lumped_fp32(gpu_ctx):
    df, dc, of, oc = levels(5)
    k = np.exp(field(oc, 1, 18))
    x = field(oc, 1, 19)
    out = vec(dc, 1, dtype=tng.F32)
    tng.KMassLumped(dc, vec(dc, 1, k, tng.F32)).apply(vec(dc, 1, x, tng.F32), out)
    assert rel_err(out.download(), O.kmass_lumped(oc, k, x)) < 2e-6


# ------------------------------------------------------------------ vector kernels
@pytest.mark.parametrize("dtype", [tng.F64, tng.F32])
@pytest.mark.parametrize("ncomp", [1, 3])
def test_vector_ops(dtype, ncomp, gpu_ctx):
    df, dc, of, oc = levels(4, 1, 1)
    tol = 1e-15 if dtype == tng.F64 else 1e-6
    a, b, c = field(of, ncomp, 20), field(of, ncomp, 21), field(of, ncomp, 22)
    va, vb, vc = vec(df, ncomp, a, dtype), vec(df, ncomp, b, dtype), vec(df, ncomp, c, dtype)
    ft = np.float64 if dtype == tng.F64 else np.float32
    a, b, c = a.astype(ft).astype(np.float64), b.astype(ft).astype(np.float64), c.astype(ft).astype(np.float64)
    y = vec(df, ncomp, dtype=dtype)
    for coef, xs, c0, exp in [([], [], 1.5, np.full_like(a, 1.5)),
                              ([2.0], [va], 0.0, 2 * a),
                              ([2.0, -0.5], [va, vb], 0.25, 2 * a - 0.5 * b + 0.25),
                              ([1.0, 2.0, 3.0], [va, vb, vc], -1.0, a + 2 * b + 3 * c - 1)]:
        tng.lincomb(y, coef, xs, c0)
        assert rel_err(y.download(), exp) < tol
    # aliasing: y = 2 y + a
    y.upload(b)
    tng.lincomb(y, [2.0, 1.0], [y, va])
    assert rel_err(y.download(), 2 * b + a) < tol
    # dot over OWNED nodes, norms, max |.|
    ref = of.masked_dot(a, b)
    assert abs(tng.dot(va, vb) - ref) <= (1e-13 if dtype == tng.F64 else 1e-5) * np.sqrt(
        of.masked_dot(a, a) * of.masked_dot(b, b))
    assert abs(tng.norm_2(va) - np.sqrt(of.masked_dot(a, a))) <= tol * 10 * np.sqrt(of.masked_dot(a, a))
    assert tng.norm_inf(vb) == np.abs(b).max()
    # scale and invert
    y.upload(b)
    tng.scale_in_place(y, va)
    assert rel_err(y.download(), a * b) < tol
    y.upload(c + 3.0)
    tng.invert_entries(y)
    assert rel_err(y.download(), 1.0 / (c.astype(ft) + ft(3.0)).astype(np.float64)) < tol
    # NaN / inf detection
    assert not tng.has_nan_or_inf(va)
    bad = a.copy()
    bad.reshape(-1)[12345] = np.inf
    y.upload(bad)
    assert tng.has_nan_or_inf(y)
    bad.reshape(-1)[12345] = np.nan
    y.upload(bad)
    assert tng.has_nan_or_inf(y)


def test_dot_deterministic(gpu_ctx):
    df, dc, of, oc = levels(5)
    a, b = field(of, 3, 23), field(of, 3, 24)
    va, vb = vec(df, 3, a), vec(df, 3, b)
    vals = {tng.dot(va, vb) for _ in range(5)}
    assert len(vals) == 1


def test_operator_argument_checks(gpu_ctx):
    df, dc, of, oc = levels(3)
    op = tng.DivergenceKerngen(df, dc, (D, D))
    with pytest.raises(tng.TngError):
        op.apply(vec(df, 1), vec(dc, 1))  # velocity must have 3 components
    with pytest.raises(tng.TngError):
        op.apply(vec(dc, 3), vec(dc, 1))  # wrong mesh
    with pytest.raises(tng.TngError):
        op.apply(vec(df, 3, dtype=tng.F32), vec(dc, 1))  # mixed dtypes
    d2, _, _, _ = levels(4)
    with pytest.raises(tng.TngError):
        tng.DivergenceKerngen(d2, dc, (D, D))  # not one refinement apart


def test_wrapped_vectors(gpu_ctx):
    """tng_vec_wrap views of existing device arrays behave like owned vectors
    (separate component arrays, as a Kokkos Grid4DDataVec holds them)."""
    df, dc, of, oc = levels(4)
    u, p = field(of, 3, 27), field(oc, 1, 28)
    k = np.exp(field(of, 1, 29))
    comps = [vec(df, 1, u[c]) for c in range(3)]  # three separate allocations
    outs = [vec(df, 1) for _ in range(3)]
    src = tng.VectorQ1Vec.wrap(df, tng.F64, [v.device_ptr() for v in comps], keepalive=comps)
    dst = tng.VectorQ1Vec.wrap(df, tng.F64, [v.device_ptr() for v in outs], keepalive=outs)
    kv = vec(df, 1, k)
    kw = tng.VectorQ1Scalar.wrap(df, tng.F64, [kv.device_ptr()], keepalive=kv)
    tng.EpsilonDivDivKerngen(df, kw, (D, F)).apply(src, dst)
    got = np.stack([o.download() for o in outs])
    assert rel_err(got, of.epsdivdiv(k, u, D, F, path="fast")) < 1e-12
    # BLAS and transfers on wrapped vectors
    tng.lincomb(dst, [2.0, -1.0], [src, dst], 0.5)
    assert rel_err(np.stack([o.download() for o in outs]), 2 * u - got + 0.5) < 1e-15
    pw = vec(dc, 1, p)
    gw = tng.VectorQ1Vec.wrap(df, tng.F64, [v.device_ptr() for v in outs], keepalive=outs)
    tng.GradientKerngen(df, dc, (D, F)).apply(tng.VectorQ1Scalar.wrap(dc, tng.F64, [pw.device_ptr()], keepalive=pw), gw)
    assert rel_err(np.stack([o.download() for o in outs]), O.gradient(of, oc, p, D, F)) < 1e-12


@pytest.mark.parametrize("L,ls,rs", [(3, 0, 0), (4, 1, 1)])
def test_vector_mass(L, ls, rs, gpu_ctx):
    df, dc, of, oc = levels(L, ls, rs)
    u = field(of, 3, 30)
    out = vec(df, 3)
    tng.VectorMass(df).apply(vec(df, 3, u), out)
    ref = O.vector_mass(of, u)
    assert rel_err(out.download(), ref) < 1e-12
    out32 = vec(df, 3, dtype=tng.F32)
    tng.VectorMass(df).apply(vec(df, 3, u, tng.F32), out32)
    assert rel_err(out32.download(), ref) < 2e-6


def test_buoyancy_rhs(gpu_ctx):
    """StokesContext::solve's RHS: f = M (Ra T x/|x|), CMB rows zeroed (Dirichlet,
    first boundary only), free-slip surface rows projected."""
    S = tng.StokesSolver(1, 3, 0.55, 1.0, (D, F))
    of = O.Mesh(3)
    T = field(of, 1, 31)
    fu = vec(S.domain_velocity, 3)
    S.buoyancy_rhs(vec(S.domain_velocity, 1, T), 1e5, fu)
    n = of.coords / np.linalg.norm(of.coords, axis=-1, keepdims=True)
    u = 1e5 * T[None] * np.moveaxis(n, -1, 0)[..., None]
    ref = O.vector_mass(of, u)
    ref[..., 0] = 0.0
    nn = np.moveaxis(n, -1, 0)
    d = (ref[..., -1] * nn).sum(0)
    ref[..., -1] -= d * nn
    assert rel_err(fu.download(), ref) < 1e-12
