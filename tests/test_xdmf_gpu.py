"""XDMF output and checkpoints (io/xdmf.hpp:148-1441) on the GPU against the
oracle's restatement of the format (oracle/xdmf.py, pinned in
test_oracle_xdmf_pins.py): every written file byte-identical, checkpoint
round trips exact (or exactly the FP32 rounding where a field is written in
4 bytes), the reference's error cases reported as errors."""
import os

import numpy as np
import pytest

import oracle as O
import tng
import xdmf as X
from conftest import consistent

pytestmark = pytest.mark.gpu


def field(om, ncomp, seed):
    r = np.random.default_rng(seed)
    return consistent(om, r.uniform(-1, 1, ((ncomp,) if ncomp > 1 else ()) + om.shape))


def vec(dom, ncomp, data=None, dtype=tng.F64):
    v = (tng.VectorQ1Vec if ncomp == 3 else tng.VectorQ1Scalar)(dom, dtype)
    if data is not None:
        v.upload(data)
    return v


def meshes(lat, rad, ls, rs):
    return (tng.DistributedDomain.create_uniform(lat, rad, 0.55, 1.0, ls, rs), O.Mesh(lat, rad, 0.55, 1.0, ls, rs))


@pytest.mark.parametrize("lat,rad,ls,rs,pb,cb", [(2, 2, 0, 0, 4, 8), (3, 3, 1, 1, 8, 4), (2, 7, 0, 0, 4, 8)])
def test_xdmf_files_match_oracle(lat, rad, ls, rs, pb, cb, tmp_path, gpu_ctx):
    dom, om = meshes(lat, rad, ls, rs)
    u, eta = field(om, 3, 1), field(om, 1, 2)
    t32, v32 = field(om, 1, 3).astype(np.float32), field(om, 3, 4).astype(np.float32)
    fields = [("velocity", vec(dom, 3, u), u, 3, "f64", 4), ("eta", vec(dom, 1, eta), eta, 1, "f64", 8),
              ("temperature", vec(dom, 1, t32, tng.F32), t32, 1, "f32", 4),
              ("v32", vec(dom, 3, v32, tng.F32), v32, 3, "f32", 8)]
    w = tng.XDMFOutput(dom, str(tmp_path / "out"), pb, cb)
    for label, v, _, _, _, ob in fields:
        w.add(v, label, ob)
    items = [{"label": lb, "data": d, "ncomp": nc, "dtype": dt, "bytes": ob} for lb, _, d, nc, dt, ob in fields]
    w.write()
    expected = X.expected_files(om, items, 0, True, pb, cb)
    # second step: new values, same fields; the mesh files are not written again
    u2 = field(om, 3, 5)
    fields[0][1].upload(u2)
    items[0]["data"] = u2
    w.write()
    expected.update(X.expected_files(om, items, 1, False, pb, cb))
    out = tmp_path / "out"
    assert sorted(os.listdir(out)) == sorted(expected)
    for name, b in expected.items():
        assert (out / name).read_bytes() == b, name
    w.close()


def test_checkpoint_round_trip(tmp_path, gpu_ctx):
    dom, om = meshes(3, 3, 1, 1)
    u, eta = field(om, 3, 6), field(om, 1, 7)
    t32 = field(om, 1, 8).astype(np.float32)
    w = tng.XDMFOutput(dom, str(tmp_path))
    vu, veta, vt = vec(dom, 3, u), vec(dom, 1, eta), vec(dom, 1, t32, tng.F32)
    w.add(vu, "u", 8)
    w.add(veta, "eta", 4)
    w.add(vt, "T", 8)
    w.set_write_counter(5)
    w.write()
    vu.upload(2.0 * u)
    w.write()  # step 6
    for step, scale in ((5, 1.0), (6, 2.0)):
        r = vec(dom, 3)
        tng.read_xdmf_checkpoint(str(tmp_path), "u", step, r)
        assert np.array_equal(r.download(), scale * u)
    r = vec(dom, 1)
    tng.read_xdmf_checkpoint(str(tmp_path), "eta", 5, r)
    assert np.array_equal(r.download(), eta.astype(np.float32).astype(np.float64))
    r32 = vec(dom, 1, dtype=tng.F32)
    tng.read_xdmf_checkpoint(str(tmp_path), "T", 6, r32)
    assert np.array_equal(r32.download(), t32)
    r32v = vec(dom, 3, dtype=tng.F32)
    tng.read_xdmf_checkpoint(str(tmp_path), "u", 5, r32v)
    assert np.array_equal(r32v.download(), u.astype(np.float32))
    w.close()


def test_checkpoint_subdomain_lookup(tmp_path, gpu_ctx):
    """The reader places every subdomain by its global id, not by file position."""
    dom, om = meshes(2, 3, 1, 1)
    eta = field(om, 1, 9)
    w = tng.XDMFOutput(dom, str(tmp_path))
    w.add(vec(dom, 1, eta), "eta", 8)
    w.write()
    # reverse the subdomain order of the file and of the id list in the metadata
    meta = bytearray((tmp_path / "checkpoint_metadata.bin").read_bytes())
    n = om.nsd
    ids = np.frombuffer(bytes(meta[-8 * n:]), dtype="<i8")
    meta[-8 * n:] = ids[::-1].tobytes()
    (tmp_path / "checkpoint_metadata.bin").write_bytes(bytes(meta))
    data = np.frombuffer((tmp_path / "eta_0.bin").read_bytes(), dtype="<f8").reshape(n, -1)
    (tmp_path / "eta_0.bin").write_bytes(data[::-1].tobytes())
    r = vec(dom, 1)
    tng.read_xdmf_checkpoint(str(tmp_path), "eta", 0, r)
    assert np.array_equal(r.download(), eta)
    w.close()


def test_xdmf_errors(tmp_path, gpu_ctx):
    dom, om = meshes(2, 2, 0, 0)
    other, _ = meshes(3, 3, 0, 0)
    v = vec(dom, 1, field(om, 1, 10))
    w = tng.XDMFOutput(dom, str(tmp_path))
    w.add(v, "eta", 8)
    with pytest.raises(tng.TngError):
        w.add(vec(dom, 1), "eta", 8)  # label taken
    with pytest.raises(tng.TngError):
        w.add(vec(dom, 1), "x", 2)  # precision
    with pytest.raises(tng.TngError):
        w.add(vec(other, 1), "y", 8)  # other mesh
    w.write()
    with pytest.raises(tng.TngError):
        w.add(vec(dom, 1), "late", 8)  # after write
    cases = [("eta", 1, vec(dom, 1)),      # missing step
             ("nope", 0, vec(dom, 1)),     # missing label
             ("eta", 0, vec(dom, 3)),      # component count
             ("eta", 0, vec(other, 1))]    # extents
    for label, step, out in cases:
        with pytest.raises(tng.TngError) as e:
            tng.read_xdmf_checkpoint(str(tmp_path), label, step, out)
        assert e.value.status == tng.IO_ERROR, (label, step)
    with pytest.raises(tng.TngError) as e:
        tng.read_xdmf_checkpoint(str(tmp_path / "missing"), "eta", 0, vec(dom, 1))
    assert e.value.status == tng.IO_ERROR
    w.close()


def test_writer_outlives_its_mesh(tmp_path, gpu_ctx):
    """Destroying a writer whose mesh is already gone (finalisers in any order) is safe."""
    import ctypes as C
    dom, om = meshes(2, 2, 0, 0)
    h = C.c_void_p()
    dom.ctx.check(tng.load_library().tng_xdmf_create(dom.h, str(tmp_path).encode(), 4, 8, C.byref(h)))
    tng.load_library().tng_mesh_destroy(dom.h)
    dom.h = None
    assert tng.load_library().tng_xdmf_destroy(h) == 0


def test_checkpoint_metadata_rejects_malformed_fields(tmp_path, gpu_ctx):
    """Negative or inconsistent fields of the checkpoint metadata are refused
    before any allocation sized from them."""
    import struct

    dom, om = meshes(2, 2, 0, 0)
    w = tng.XDMFOutput(dom, str(tmp_path))
    w.add(vec(dom, 1, field(om, 1, 12)), "eta", 8)
    w.write()
    w.close()
    good = (tmp_path / "checkpoint_metadata.bin").read_bytes()
    version, S, Sr, sx, sy, sr = struct.unpack_from("<6i", good, 0)
    off_files = 24 + 8 * (Sr * (sr - 1) + 1) + (4 if version > 0 else 0)
    bad = []
    for off, val in [(4, -1), (8, 0), (20, 1), (4, 1 << 30), (off_files + 4, -5)]:
        b = bytearray(good)
        struct.pack_into("<i", b, off, val)
        bad.append(bytes(b))
    bad.append(good[: len(good) // 2])  # truncated
    for b in bad:
        (tmp_path / "checkpoint_metadata.bin").write_bytes(b)
        with pytest.raises(tng.TngError) as e:
            tng.read_xdmf_checkpoint(str(tmp_path), "eta", 0, vec(dom, 1))
        assert e.value.status == tng.IO_ERROR
    (tmp_path / "checkpoint_metadata.bin").write_bytes(good)
    r = vec(dom, 1)
    tng.read_xdmf_checkpoint(str(tmp_path), "eta", 0, r)
