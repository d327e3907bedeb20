"""The allocation-free EpsilonDivDiv apply (tng_epsdivdiv_apply_noalloc)
through the C ABI against tng_op_apply on the same operator and inputs: the
same kernels, so the results are bit-identical.  Two calls in a row, the
second on a smaller mesh (the shared scratch keeps the first call's
contents), free slip on both surfaces (penalty moments in the scratch) and
128 radial layers (partial sums between radial parts in the scratch); then a
mesh whose partial sums exceed the scratch is rejected."""
import ctypes

import numpy as np
import pytest

import tng
from tng import inputs

pytestmark = pytest.mark.gpu

F, D = tng.FREESLIP, tng.DIRICHLET


def _case(L, Lr, bcs, seed):
    dom = tng.DistributedDomain.create_uniform(L, Lr, 0.55, 1.0)
    w = inputs.Workload(dom.coords(), dom.radii(), L, Lr)
    kv = tng.VectorQ1Scalar(dom).upload(w.eta_layered(seed))
    op = tng.EpsilonDivDivKerngen(dom, kv, bcs)
    src = tng.VectorQ1Vec(dom).upload(w.velocity(seed + 1))
    return dom, kv, op, src


def _noalloc(op, src, dst):
    lib = tng.load_library()
    st = lib.tng_epsdivdiv_apply_noalloc(op.h, src.h, dst.h)
    return st


@pytest.mark.parametrize("bcs", [(F, F), (D, F)])
def test_noalloc_apply_matches_apply_twice(bcs, gpu_ctx):
    for L, Lr, seed in [(3, 7, 201), (2, 7, 203)]:  # the second call on a smaller mesh
        dom, kv, op, src = _case(L, Lr, bcs, seed)
        ref, got = tng.VectorQ1Vec(dom), tng.VectorQ1Vec(dom)
        op.apply(src, ref)
        got.fill(np.nan)
        assert _noalloc(op, src, got) == 0, tng.load_library().tng_last_error(dom.ctx.h)
        dom.ctx.synchronize()
        a, b = ref.download(), got.download()
        assert np.array_equal(a, b), (L, Lr, float(np.nanmax(np.abs(a - b))))
        # Add mode and 64 layers (no radial parts) through the same entry point
        op.set_operator_apply_and_communication_modes(tng.ADD, tng.COMMUNICATE_ADDITIVELY)
        op.apply(src, ref)
        assert _noalloc(op, src, got) == 0
        assert np.array_equal(ref.download(), got.download())


def test_noalloc_apply_on_a_fresh_operator(gpu_ctx):
    """The first apply of an operator through the allocation-free entry point
    (its boundary rows computed in the call) equals an ordinary apply."""
    dom, kv, op, src = _case(3, 6, (F, D), 207)
    got = tng.VectorQ1Vec(dom)
    assert _noalloc(op, src, got) == 0, tng.load_library().tng_last_error(dom.ctx.h)
    op2 = tng.EpsilonDivDivKerngen(dom, kv, (F, D))
    ref = tng.VectorQ1Vec(dom)
    op2.apply(src, ref)
    assert np.array_equal(ref.download(), got.download())


def test_noalloc_apply_single_part_and_fp32(gpu_ctx):
    for dt in (tng.F64, tng.F32):
        dom = tng.DistributedDomain.create_uniform(3, 3, 0.55, 1.0)
        w = inputs.Workload(dom.coords(), dom.radii(), 3, 3)
        kv = tng.VectorQ1Scalar(dom, dt).upload(w.eta_layered(205))
        op = tng.EpsilonDivDivKerngen(dom, kv, (F, F))
        src = tng.VectorQ1Vec(dom, dt).upload(w.velocity(206))
        ref, got = tng.VectorQ1Vec(dom, dt), tng.VectorQ1Vec(dom, dt)
        op.apply(src, ref)
        assert _noalloc(op, src, got) == 0
        assert np.array_equal(ref.download(), got.download())


def test_noalloc_apply_rejects_oversized_scratch(gpu_ctx):
    """Level 6 with 512 layers: eight radial parts need 14.2 MB of partial sums."""
    dom = tng.DistributedDomain.create_uniform(6, 9, 0.55, 1.0)
    kv = tng.VectorQ1Scalar(dom)
    kv.fill(1.0)
    op = tng.EpsilonDivDivKerngen(dom, kv, (D, D))
    src, dst = tng.VectorQ1Vec(dom), tng.VectorQ1Vec(dom)
    st = _noalloc(op, src, dst)
    assert st == 1  # TNG_INVALID_ARGUMENT
    msg = tng.load_library().tng_last_error(dom.ctx.h).decode()
    assert "scratch" in msg, msg
