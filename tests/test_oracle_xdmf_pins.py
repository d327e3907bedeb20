"""Pins of the oracle's restatement of the XDMF / checkpoint format
(io/xdmf.hpp:74-147, 578-651, 898-1040; util/xml.hpp) against what the
format documentation and the mesh fix.  No GPU needed.

- Field files follow the documented loop nest (subdomain, r, y, x, component),
  checked by explicit loops on a small mesh.
- geometry.bin holds points whose norm is the node's shell radius and whose
  direction is the unit-sphere node coordinate.
- topology.bin: wedge 0 of a cell is the triangle (x, y), (x+1, y), (x, y+1)
  and wedge 1 the triangle (x+1, y+1), (x, y+1), (x+1, y), each repeated one
  radial level up (the wedge node order of the operators).
- checkpoint_metadata.bin parses with the documented field sequence, and the
  global subdomain ids decode to the subdomains' (diamond, x, y, r).
- step_<n>.xmf parses as XML with the documented structure.
"""
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle as O
import xdmf as X


@pytest.mark.parametrize("ncomp,out", [(1, 8), (1, 4), (3, 8), (3, 4)])
def test_field_bytes_loop_nest(ncomp, out):
    m = O.Mesh(1, 2)
    r = np.random.default_rng(1)
    f = r.standard_normal(((ncomp,) if ncomp > 1 else ()) + m.shape)
    vals = []
    for sd in range(m.nsd):
        for rr in range(m.nr):
            for y in range(m.nx):
                for x in range(m.nx):
                    for d in range(ncomp):
                        vals.append(f[sd, x, y, rr] if ncomp == 1 else f[d, sd, x, y, rr])
    want = np.array(vals).astype(np.float64 if out == 8 else np.float32).tobytes()
    assert X.field_bytes(f, out) == want


@pytest.mark.parametrize("lat,rad,ls,rs", [(2, 2, 0, 0), (2, 3, 1, 1)])
def test_geometry_radius_and_direction(lat, rad, ls, rs):
    m = O.Mesh(lat, rad, 0.55, 1.0, ls, rs)
    g = np.frombuffer(X.geometry_bytes(m, 8), dtype=np.float64).reshape(m.nsd, m.nr, m.nx, m.nx, 3)
    norm = np.linalg.norm(g, axis=-1)
    assert np.allclose(norm, m.radii[:, :, None, None], rtol=1e-14, atol=0)
    unit = g / norm[..., None]
    assert np.allclose(unit, m.coords.transpose(0, 2, 1, 3)[:, None], rtol=0, atol=1e-14)
    g4 = np.frombuffer(X.geometry_bytes(m, 4), dtype=np.float32)
    assert np.array_equal(g4, g.ravel().astype(np.float32))


@pytest.mark.parametrize("ib", [4, 8])
def test_topology_wedges(ib):
    m = O.Mesh(2, 2)
    n, nr = m.nx, m.nr
    t = np.frombuffer(X.topology_bytes(m, ib), dtype=np.int32 if ib == 4 else np.int64)
    t = t.reshape(m.nsd, nr - 1, n - 1, n - 1, 2, 6)

    def node(sd, r, y, x):
        return ((sd * nr + r) * n + y) * n + x

    for sd in (0, 7):
        for r in range(nr - 1):
            for y in range(n - 1):
                for x in range(n - 1):
                    w0 = [node(sd, r, y, x), node(sd, r, y, x + 1), node(sd, r, y + 1, x)]
                    w1 = [node(sd, r, y + 1, x + 1), node(sd, r, y + 1, x), node(sd, r, y, x + 1)]
                    up = n * n
                    assert list(t[sd, r, y, x, 0]) == w0 + [v + up for v in w0]
                    assert list(t[sd, r, y, x, 1]) == w1 + [v + up for v in w1]
    assert t.max() < m.nalloc


def test_metadata_parses_as_documented():
    m = O.Mesh(2, 3, 0.55, 1.0, 1, 1)
    items = [{"label": "u", "ncomp": 3, "dtype": "f64", "bytes": 4},
             {"label": "eta", "ncomp": 1, "dtype": "f64", "bytes": 8},
             {"label": "T32", "ncomp": 1, "dtype": "f32", "bytes": 4}]
    b = X.metadata_bytes(m, items, 4)
    pos = 0

    def take(fmt):
        nonlocal pos
        v = struct.unpack_from("<" + fmt, b, pos)
        pos += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]

    assert take("i") == 1
    S, Sr, sx, sy, sr = take("iiiii")
    assert (S, Sr, sx, sy, sr) == (2, 2, m.nx, m.nx, m.nr)
    radii = np.array(take("d" * (Sr * (sr - 1) + 1)))
    assert np.all(np.diff(radii) > 0) and radii[0] == 0.55 and radii[-1] == 1.0
    assert take("i") == 4
    nf = take("i")
    names = []
    for _ in range(nf):
        ln = take("i")
        names.append(b[pos:pos + ln].decode())
        pos += ln
        assert take("i") == 2
        take("ii")
    # scalar float, scalar double, vector float, vector double
    assert names == ["T32", "eta", "u"]
    assert take("i") == 0
    ids = [take("q") for _ in range(10 * S * S * Sr)]
    assert pos == len(b)
    info = m.subdomain_info
    for sd, gid in enumerate(ids):
        dec = (gid >> 57, gid & ((1 << 19) - 1), (gid >> 19) & ((1 << 19) - 1), (gid >> 38) & ((1 << 19) - 1))
        assert dec == tuple(info[sd])


def test_xmf_structure():
    m = O.Mesh(2, 2)
    items = [{"label": "velocity", "ncomp": 3, "dtype": "f64", "bytes": 4},
             {"label": "p&q", "ncomp": 1, "dtype": "f64", "bytes": 8}]
    text = X.xmf_text(m, items, 3, 4, 8)
    assert text.startswith('<?xml version="1.0" ?>\n<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>\n<Xdmf Version="2.0">\n')
    root = ET.fromstring(text.split("\n", 2)[2])
    grid = root.find("Domain/Grid")
    assert grid.attrib == {"Name": "Grid", "GridType": "Uniform"}
    geo = grid.find("Geometry/DataItem")
    assert geo.text == "geometry.bin" and geo.attrib["Dimensions"] == f"{m.nalloc} 3"
    topo = grid.find("Topology")
    ne = m.nsd * (m.nx - 1) ** 2 * (m.nr - 1) * 2
    assert topo.attrib == {"Type": "Wedge", "NumberOfElements": str(ne)}
    assert topo.find("DataItem").attrib["Dimensions"] == str(6 * ne)
    attrs = grid.findall("Attribute")
    assert [a.attrib["Name"] for a in attrs] == ["p&q", "velocity"]
    assert attrs[0].find("DataItem").text == "p&q_3.bin" and attrs[0].find("DataItem").attrib["Precision"] == "8"
    assert attrs[1].attrib["AttributeType"] == "Vector"
    # attributes in name order, as util/xml.hpp's std::map writes them
    assert '<DataItem DataType="Float" Dimensions="' in text and "p&amp;q_3.bin" in text


def test_python_metadata_reader(tmp_path):
    """tng.read_xdmf_checkpoint_metadata (host only) parses what the format writes."""
    import tng
    m = O.Mesh(2, 3, 0.55, 1.0, 1, 1)
    items = [{"label": "u", "ncomp": 3, "dtype": "f64", "bytes": 4},
             {"label": "eta", "ncomp": 1, "dtype": "f32", "bytes": 8}]
    (tmp_path / "checkpoint_metadata.bin").write_bytes(X.metadata_bytes(m, items, 8))
    md = tng.read_xdmf_checkpoint_metadata(str(tmp_path))
    assert md["version"] == 1 and md["grid_scalar_bytes"] == 8
    assert (md["size_x"], md["size_y"], md["size_r"]) == (m.nx, m.nx, m.nr)
    assert [f["grid_name_string"] for f in md["grid_data_files"]] == ["eta", "u"]
    assert md["grid_data_files"][1]["vec_dim"] == 3 and md["grid_data_files"][0]["scalar_bytes"] == 8
    assert md["checkpoint_ordering_0_global_subdomain_ids"] == [X.global_subdomain_id(*r) for r in m.subdomain_info]
    (tmp_path / "checkpoint_metadata.bin").write_bytes(X.metadata_bytes(m, items, 8)[:-3])
    with pytest.raises(tng.TngError):
        tng.read_xdmf_checkpoint_metadata(str(tmp_path))
    with pytest.raises(tng.TngError):
        tng.read_xdmf_checkpoint_metadata(str(tmp_path / "nope"))
