"""Parity of the HIP EpsilonDivDiv apply (through the C ABI) with the CPU
oracle, which restates the reference's local-matrix and fast paths
(fe/wedge/operators/shell/epsilon_divdiv_kerngen.hpp:577-792, 1013-1242,
1245-1607, 1638-2198).  FP64 bar: 1e-12 relative (||diff||_2 / ||ref||_2).
FP32 bar: 2e-6 relative."""
import numpy as np
import pytest

import oracle as O
import tng
from conftest import consistent, rel_err
from tng import inputs

pytestmark = pytest.mark.gpu

N, D, F = tng.NEUMANN, tng.DIRICHLET, tng.FREESLIP

_cache = {}


def setup(L, Lr=None, ls=0, rs=0, rmin=0.55, rmax=1.0):
    Lr = L if Lr is None else Lr
    key = (L, Lr, ls, rs, rmin, rmax)
    if key not in _cache:
        dom = tng.DistributedDomain.create_uniform(L, Lr, rmin, rmax, ls, rs)
        om = O.Mesh(L, Lr, rmin, rmax, ls, rs)
        w = inputs.Workload(dom.coords(), dom.radii(), L, Lr, ls, rs)
        _cache[key] = (dom, om, w)
    return _cache[key]


def gpu_apply(dom, k, src, bcs, diagonal=False, apply_mode=tng.REPLACE, comm=tng.COMMUNICATE_ADDITIVELY,
              eps=None, dst0=None, dtype=tng.F64):
    kv = tng.VectorQ1Scalar(dom, dtype).upload(k)
    op = tng.EpsilonDivDivKerngen(dom, kv, bcs, diagonal)
    op.set_operator_apply_and_communication_modes(apply_mode, comm)
    if eps is not None:
        op.set_penalty_epsilon(eps)
    s = tng.VectorQ1Vec(dom, dtype).upload(src)
    d = tng.VectorQ1Vec(dom, dtype)
    if dst0 is not None:
        d.upload(dst0)
    op.apply(s, d)
    return d.download().astype(np.float64)


def test_config1_level4_dirichlet(gpu_ctx):
    """Config 1: L4 shell r in [0.55, 1], eta = exp(z), u ~ U[-1, 1], Dirichlet
    on both surfaces; compared with the reference's local-matrix path."""
    dom, om, w = setup(4)
    k = w.eta_lognormal(102)
    u = w.velocity(101)
    g = gpu_apply(dom, k, u, (D, D))
    ref = om.epsdivdiv(k, u, D, D, path="slow")
    own = om.ownership_mask == 1
    assert rel_err(g, ref, own) < 1e-12
    assert rel_err(g, ref) < 1e-12


BCS = [(N, N), (D, D), (F, F), (F, D), (D, F), (N, F), (F, N), (D, N)]


@pytest.mark.parametrize("bcs", BCS)
@pytest.mark.parametrize("diag", [False, True])
def test_parity_bcs(gpu_ctx, bcs, diag):
    dom, om, w = setup(3)
    k = w.eta_layered(102)
    u = w.velocity(101)
    g = gpu_apply(dom, k, u, bcs, diagonal=diag, eps=1.0)
    ref = om.epsdivdiv(k, u, *bcs, diagonal=diag, path="fast", penalty_eps=1.0)
    assert rel_err(g, ref) < 1e-12


@pytest.mark.parametrize("L,Lr,ls,rs", [(1, 1, 0, 0), (2, 2, 0, 0), (2, 4, 0, 0), (3, 3, 1, 0), (3, 3, 0, 1),
                                        (3, 3, 1, 1), (2, 5, 0, 0), (2, 6, 0, 0), (2, 7, 0, 0), (2, 7, 0, 1),
                                        (3, 6, 1, 0), (3, 7, 1, 0), (4, 6, 2, 0), (2, 8, 0, 1), (1, 9, 0, 1)])
@pytest.mark.parametrize("bcs", [(D, D), (F, F)])
def test_parity_shapes(gpu_ctx, L, Lr, ls, rs, bcs):
    """Every lane mapping (4..64 radial lanes, radial parts for 128 and 256
    layers) and subdomain split, radial parts within radially split
    subdomains included."""
    dom, om, w = setup(L, Lr, ls, rs)
    k = w.eta_lognormal(102)
    u = w.velocity(101)
    g = gpu_apply(dom, k, u, bcs, eps=1.0)
    ref = om.epsdivdiv(k, u, *bcs, path="fast", penalty_eps=1.0)
    assert rel_err(g, ref) < 1e-12


@pytest.mark.parametrize("L,Lr,env", [(2, 7, {"apply_radial_chunks": 1}), (3, 6, {"apply_waves": 4}),
                                       (1, 8, {"apply_radial_chunks": 1}), (1, 8, {"apply_radial_split": 0}),
                                       (1, 8, {}), (2, 8, {}), (2, 7, {"apply_radial_split": 1}), (1, 9, {}),
                                       (3, 7, {}), (3, 6, {})])
@pytest.mark.parametrize("bcs", [(D, D), (F, F), (N, N), (D, F), (N, F), (F, D)])
def test_parity_launch_variants(gpu_ctx, L, Lr, env, bcs):
    """Launch shapes chosen by layer count or tuning variable: 128, 256 and
    512 layers as radial parts of 64 cells swept separately; 128 and 256 as
    radial chunks of one wave each in one workgroup, or as two or four cells
    per lane; 4-wave workgroups at 64 layers; every boundary-condition mix of
    the one-column kernels."""
    dom, om, w = setup(L, Lr)
    k = w.eta_lognormal(102)
    u = w.velocity(101)
    for add in (False, True):
        d0 = w.velocity(103)
        with tng.default_context().tuning(**env):
            g = gpu_apply(dom, k, u, bcs, eps=1.0, apply_mode=tng.ADD if add else tng.REPLACE, dst0=d0)
        ref = om.epsdivdiv(k, u, *bcs, path="fast", penalty_eps=1.0, apply_mode=tng.ADD if add else tng.REPLACE,
                           dst=d0)
        assert rel_err(g, ref) < 1e-12


@pytest.mark.parametrize("Lr", [7, 8])
@pytest.mark.parametrize("bcs", [(D, D), (F, F), (F, D)])
@pytest.mark.parametrize("dtype", [tng.F64, tng.F32])
def test_radial_parts_match_radial_chunks(gpu_ctx, Lr, bcs, dtype):
    """128 and 256 layers: the separately swept radial parts plus the
    interface kernel give the same bits as radial chunks in one workgroup
    (same sums in the same order), in both apply modes."""
    dom, om, w = setup(2, Lr)
    k = w.eta_lognormal(102)
    u = w.velocity(101)
    d0 = w.velocity(103)
    out = {}
    for split in (1, 0):
        with tng.default_context().tuning(apply_radial_split=split):
            out[split] = [gpu_apply(dom, k, u, bcs, eps=1.0, dtype=dtype, apply_mode=m, dst0=d0)
                          for m in (tng.REPLACE, tng.ADD)]
    for a, b in zip(out[1], out[0]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("dtype,tol", [(tng.F64, 1e-12), (tng.F32, 2e-6)])
def test_penalty_moment_unaligned_source(gpu_ctx, dtype, tol):
    """Free slip on both surfaces (rotation penalty on) with wrapped source
    components that start one value past an aligned address."""
    dom, om, w = setup(3)
    big, _, _ = setup(4)  # raw storage with room for one extra value
    k = w.eta_layered(102)
    u = w.velocity(101)
    n = dom.num_alloc
    store = []
    for c in range(3):
        raw = np.zeros(big.num_alloc)
        raw[1:1 + n] = u[c].ravel()
        store.append(tng.VectorQ1Scalar(big, dtype).upload(raw.reshape(big.shape)))
    esize = 8 if dtype == tng.F64 else 4
    src = tng.VectorQ1Vec.wrap(dom, dtype, [v.device_ptr() + esize for v in store], keepalive=store)
    kv = tng.VectorQ1Scalar(dom, dtype).upload(k)
    op = tng.EpsilonDivDivKerngen(dom, kv, (F, F))
    op.set_penalty_epsilon(1.0)
    d = tng.VectorQ1Vec(dom, dtype)
    op.apply(src, d)
    ref = om.epsdivdiv(k, u, F, F, path="fast", penalty_eps=1.0)
    assert rel_err(d.download().astype(np.float64), ref) < tol
    aligned = gpu_apply(dom, k, u, (F, F), eps=1.0, dtype=dtype)
    assert rel_err(d.download().astype(np.float64), aligned) < 50 * np.finfo(np.float64 if esize == 8 else np.float32).eps


@pytest.mark.parametrize("mode", [(tng.REPLACE, tng.SKIP_COMMUNICATION), (tng.ADD, tng.SKIP_COMMUNICATION),
                                  (tng.ADD, tng.COMMUNICATE_ADDITIVELY)])
@pytest.mark.parametrize("bcs", [(D, D), (F, F)])
def test_apply_and_communication_modes(gpu_ctx, mode, bcs):
    dom, om, w = setup(3)
    rng = np.random.default_rng(11)
    k = w.eta_lognormal(102)
    u = rng.uniform(-1, 1, (3,) + dom.shape)  # inconsistent copies on purpose
    d0 = rng.uniform(-1, 1, (3,) + dom.shape)
    g = gpu_apply(dom, k, u, bcs, apply_mode=mode[0], comm=mode[1], eps=1.0, dst0=d0)
    ref = om.epsdivdiv(k, u, *bcs, apply_mode=mode[0], comm=mode[1], penalty_eps=1.0, dst=d0, path="fast")
    assert rel_err(g, ref) < 1e-12


def test_config2_shape_level5_freeslip(gpu_ctx):
    """Config-2 coefficients (layered 1e3 jump at r = 0.85 times log-normal
    sigma 0.5), free-slip both surfaces, default penalty 1e-7, at L5."""
    dom, om, w = setup(5)
    k = w.eta_layered(102)
    u = w.velocity(101)
    g = gpu_apply(dom, k, u, (F, F))
    ref = om.epsdivdiv(k, u, F, F, path="fast")
    assert rel_err(g, ref) < 1e-12


def test_config2_level6_freeslip_timed_size(gpu_ctx):
    """The timed apply workload itself: level 6 (10 x 64x64 x 64 cells),
    free slip on both surfaces, default penalty 1e-7, layered 1e3 jump at
    r = 0.85 times log-normal(0.5), u ~ U[-1, 1] (seeds 101/102).  This runs
    the 64-lane FP64 kernel with one round of resident workgroups, each
    sweeping a share of many node rows across strip boundaries."""
    dom, om, w = setup(6)
    k = w.eta_layered(102)
    u = w.velocity(101)
    g = gpu_apply(dom, k, u, (F, F))
    ref = om.epsdivdiv(k, u, F, F, path="fast")
    assert rel_err(g, ref) < 1e-12
    own = om.ownership_mask == 1
    assert rel_err(g, ref, own) < 1e-12


@pytest.mark.parametrize("L,blocks", [(3, 5), (3, 7), (4, 3), (4, 37)])
@pytest.mark.parametrize("bcs", [(F, F), (D, D), (F, D)])
def test_balanced_shares_cross_strips(gpu_ctx, L, blocks, bcs):
    """Few workgroups: every share of node rows runs over several strip
    boundaries (and subdomain boundaries inside a strip)."""
    dom, om, w = setup(L)
    k = w.eta_layered(102)
    u = w.velocity(101)
    d0 = w.velocity(103)
    for mode in (tng.REPLACE, tng.ADD):
        with tng.default_context().tuning(apply_blocks=blocks):
            g = gpu_apply(dom, k, u, bcs, eps=1.0, apply_mode=mode, dst0=d0)
        ref = om.epsdivdiv(k, u, *bcs, path="fast", penalty_eps=1.0, apply_mode=mode, dst=d0)
        assert rel_err(g, ref) < 1e-12


# FP32 rounding is amplified by the cancellation between lateral and radial
# stencil terms, which grows with the cells' aspect ratio (lateral width over
# radial height: about 1.4 at (4, 4), 40 at (2, 7), 80 at (2, 8)).
@pytest.mark.parametrize("L,Lr,tol", [(4, 4, 2e-6), (2, 7, 2e-5), (2, 8, 4e-5)])
def test_fp32_apply(gpu_ctx, L, Lr, tol):
    dom, om, w = setup(L, Lr)
    k = w.eta_layered(102)
    u = w.velocity(101)
    for bcs in [(D, D), (F, F)]:
        g = gpu_apply(dom, k, u, bcs, dtype=tng.F32)
        ref = om.epsdivdiv(k, u, *bcs, path="fast")
        assert rel_err(g, ref) < tol


def test_deterministic(gpu_ctx):
    dom, om, w = setup(4)
    k = w.eta_layered(102)
    u = w.velocity(101)
    a = gpu_apply(dom, k, u, (F, F))
    b = gpu_apply(dom, k, u, (F, F))
    assert np.array_equal(a, b)


def test_coefficient_update(gpu_ctx):
    """k is captured by reference; boundary rows follow after
    coefficient_updated()."""
    dom, om, w = setup(3)
    k1 = w.eta_lognormal(102)
    k2 = w.eta_lognormal(103)
    u = w.velocity(101)
    kv = tng.VectorQ1Scalar(dom).upload(k1)
    op = tng.EpsilonDivDivKerngen(dom, kv, (D, F))
    s = tng.VectorQ1Vec(dom).upload(u)
    d = tng.VectorQ1Vec(dom)
    op.apply(s, d)
    assert rel_err(d.download(), om.epsdivdiv(k1, u, D, F, path="fast")) < 1e-12
    kv.upload(k2)
    op.coefficient_updated()
    op.apply(s, d)
    assert rel_err(d.download(), om.epsdivdiv(k2, u, D, F, path="fast")) < 1e-12
