"""Drop-in entry points of the operator, GCA and vector interfaces through the
C ABI, against the oracle:
  - EpsDivDivStokes (fe/wedge/operators/shell/epsilon_divdiv_stokes.hpp:18-133)
    with its block accessors (linalg/operator.hpp:117-141);
  - per-element GCACapable get/set_local_matrix (linalg/operator.hpp:80-112,
    epsilon_divdiv_kerngen.hpp:503-575);
  - norm_2, randomize and a dot left on the device (linalg/vector.hpp:7-42).
FP64 bar 1e-12 relative."""
import ctypes

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


def vec(dom, ncomp, data=None, dtype=tng.F64):
    v = (tng.VectorQ1Vec if ncomp == 3 else tng.VectorQ1Scalar)(dom, dtype)
    if data is not None:
        v.upload(data)
    return v


@pytest.mark.parametrize("bcs", [(D, D), (F, F), (D, F), (N, N)])
def test_stokes_operator_apply_and_blocks(bcs, gpu_ctx):
    df, dc = tng.DistributedDomain.create_uniform(3, 3, 0.55, 1.0), tng.DistributedDomain.create_uniform(2, 2, 0.55, 1.0)
    of, oc = O.Mesh(3), O.Mesh(2)
    k = np.exp(0.5 * field(of, 1, 5))
    u, p = field(of, 3, 6), field(oc, 1, 7)
    kv = vec(df, 1, k)
    K = tng.EpsDivDivStokes(df, dc, kv, bcs)
    ou, op_ = vec(df, 3), vec(dc, 1)
    K.apply(vec(df, 3, u), vec(dc, 1, p), ou, op_)
    ru = of.epsdivdiv(k, u, *bcs, apply_mode=O.REPLACE, comm=O.SKIP, path="fast")
    ru = O.gradient(of, oc, p, *bcs, apply_mode=O.ADD, comm=O.COMMUNICATE, dst=ru)
    rp = O.divergence(of, oc, u, *bcs)
    assert rel_err(ou.download(), ru) < 1e-12
    assert rel_err(op_.download(), rp) < 1e-12
    # the blocks, each applied alone (left in Replace + Communicate by the Stokes apply)
    a = vec(df, 3)
    K.block_11().apply(vec(df, 3, u), a)
    assert rel_err(a.download(), of.epsdivdiv(k, u, *bcs, path="fast")) < 1e-12
    g = vec(df, 3)
    K.block_12().apply(vec(dc, 1, p), g)
    assert rel_err(g.download(), O.gradient(of, oc, p, *bcs)) < 1e-12
    b = vec(dc, 1)
    K.block_21().apply(vec(df, 3, u), b)
    assert rel_err(b.download(), rp) < 1e-12
    z = vec(dc, 1, p)
    K.block_22().apply(vec(dc, 1, p), z)
    assert np.count_nonzero(z.download()) == 0
    # a block handle is released with the Stokes operator, not on its own
    lib = tng.load_library()
    assert lib.tng_op_destroy(K.block_11().h) != 0


def test_stokes_operator_diagonal(gpu_ctx):
    df, dc = tng.DistributedDomain.create_uniform(2, 2, 0.55, 1.0), tng.DistributedDomain.create_uniform(1, 1, 0.55, 1.0)
    of = O.Mesh(2)
    k = np.exp(0.5 * field(of, 1, 8))
    u = field(of, 3, 9)
    K = tng.EpsDivDivStokes(df, dc, vec(df, 1, k), (D, F), diagonal=True)
    p0 = np.arange(np.prod(dc.shape), dtype=np.float64).reshape(dc.shape)
    ou, op_ = vec(df, 3), vec(dc, 1, p0)
    K.apply(vec(df, 3, u), vec(dc, 1), ou, op_)
    assert rel_err(ou.download(), of.epsdivdiv(k, u, D, F, diagonal=True, path="fast")) < 1e-12
    assert np.array_equal(op_.download(), p0)  # pressure block untouched
    with pytest.raises(tng.TngError):
        K.block_12()


def test_local_matrix_per_element(gpu_ctx):
    dom = tng.DistributedDomain.create_uniform(2, 2, 0.55, 1.0)
    om = O.Mesh(2)
    k = np.exp(field(om, 1, 10))
    op = tng.EpsilonDivDivKerngen(dom, vec(dom, 1, k), (D, D))
    ref = O.local_matrices(om, k)
    # without storage: the assembled Neumann matrix
    for idx in [(0, 0, 0, 0, 0), (3, 1, 2, 3, 1), (9, 3, 3, 3, 1)]:
        assert rel_err(op.get_local_matrix(*idx), ref[idx]) < 1e-13
    # set needs stored mode, as the reference asserts
    with pytest.raises(tng.TngError):
        op.set_local_matrix(0, 0, 0, 0, 0, ref[0, 0, 0, 0, 0])
    op.set_stored_matrix_mode(True)
    assert np.count_nonzero(op.get_local_matrix(1, 1, 1, 1, 0)) == 0  # Full storage starts at zero
    rng = np.random.default_rng(11)
    lm = np.zeros_like(ref)
    # one element at a time into a zero store, then the stored apply against the oracle's slow path
    stored = ref * (1.0 + 0.1 * rng.uniform(size=ref.shape[:5] + (1, 1)))
    for idx in np.ndindex(ref.shape[:5]):
        op.set_local_matrix(*idx, stored[idx])
        lm[idx] = stored[idx]
    assert rel_err(op.get_local_matrix(4, 2, 1, 0, 1), stored[4, 2, 1, 0, 1]) == 0.0
    assert rel_err(op.get_local_matrices(), lm) == 0.0
    u = field(om, 3, 12)
    out = vec(dom, 3)
    op.apply(vec(dom, 3, u), out)
    assert rel_err(out.download(), O.epsdivdiv_stored(om, lm, u, D, D)) < 1e-12
    with pytest.raises(tng.TngError):
        op.get_local_matrix(0, 4, 0, 0, 0)  # cell index out of range
    op.set_stored_matrix_mode(False)
    assert rel_err(op.get_local_matrix(3, 1, 2, 3, 1), ref[3, 1, 2, 3, 1]) < 1e-13


def test_set_boundary_conditions(gpu_ctx):
    """set_boundary_conditions (epsilon_divdiv_kerngen.hpp:474-481) changes the
    rows the next apply treats; the rotation penalty stays as armed at
    creation (an operator created with Neumann rows has none)."""
    dom = tng.DistributedDomain.create_uniform(3, 3, 0.55, 1.0)
    om = O.Mesh(3)
    k = np.exp(0.5 * field(om, 1, 20))
    u = field(om, 3, 21)
    op = tng.EpsilonDivDivKerngen(dom, vec(dom, 1, k), (N, N))
    out = vec(dom, 3)
    op.apply(vec(dom, 3, u), out)
    assert rel_err(out.download(), om.epsdivdiv(k, u, N, N, path="fast")) < 1e-12
    for bcs in [(D, F), (F, D), (D, D), (F, F)]:
        op.set_boundary_conditions(bcs)
        op.apply(vec(dom, 3, u), out)
        assert rel_err(out.download(), om.epsdivdiv(k, u, *bcs, path="fast", penalty_eps=0.0)) < 1e-12
    with pytest.raises(tng.TngError):
        op.set_boundary_conditions((N, 7))


def test_clear_local_matrices(gpu_ctx):
    """clear_local_matrices drops the stored matrices: get returns the
    assembled Neumann matrices again and the apply is matrix-free."""
    dom = tng.DistributedDomain.create_uniform(2, 2, 0.55, 1.0)
    om = O.Mesh(2)
    k = np.exp(field(om, 1, 22))
    op = tng.EpsilonDivDivKerngen(dom, vec(dom, 1, k), (D, F))
    ref = O.local_matrices(om, k)
    op.set_stored_matrix_mode(True)
    op.set_local_matrix(2, 1, 0, 1, 1, 2.0 * ref[2, 1, 0, 1, 1])
    assert rel_err(op.get_local_matrix(2, 1, 0, 1, 1), 2.0 * ref[2, 1, 0, 1, 1]) == 0.0
    op.clear_local_matrices()
    assert rel_err(op.get_local_matrix(2, 1, 0, 1, 1), ref[2, 1, 0, 1, 1]) < 1e-13
    u = field(om, 3, 23)
    out = vec(dom, 3)
    op.apply(vec(dom, 3, u), out)
    assert rel_err(out.download(), om.epsdivdiv(k, u, D, F, path="fast")) < 1e-12


def test_vector_norm_randomize_device_dot(gpu_ctx):
    dom = tng.DistributedDomain.create_uniform(3, 3, 0.55, 1.0)
    om = O.Mesh(3)
    own = om.ownership_mask == 1
    x = field(om, 3, 13)
    y = field(om, 3, 14)
    xv, yv = vec(dom, 3, x), vec(dom, 3, y)
    assert abs(tng.norm_2(xv) - np.sqrt((x[:, own] ** 2).sum())) < 1e-12 * np.sqrt((x[:, own] ** 2).sum())
    # randomize: uniform [0, 1) on every copy, reproducible from the seed, seed-dependent
    r1, r2, r3 = vec(dom, 3), vec(dom, 3), vec(dom, 3)
    tng.randomize(r1, 7)
    tng.randomize(r2, 7)
    tng.randomize(r3, 8)
    a, b, c = r1.download(), r2.download(), r3.download()
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 0.0 and a.max() < 1.0 and abs(a.mean() - 0.5) < 0.01
    # dot written to device memory, read back through a vector's storage
    slot = vec(dom, 1)
    tng.dot_device(xv, yv, slot.device_ptr(0) + 8 * 5)
    got = slot.download().ravel()[5]
    assert got == tng.dot(xv, yv)
    assert abs(got - (x[:, own] * y[:, own]).sum()) < 1e-12 * np.abs(x[:, own] * y[:, own]).sum()
