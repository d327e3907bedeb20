"""libtng on the host only (no GPU): the C-ABI library loads, exports every
symbol include/*.h declares, and its host-built mesh equals the oracle's bit
for bit."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import oracle as O
import tng

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(tng_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(tng.LIB_PATH)
    names = declared_symbols()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python front end binds exactly the declared entry points
    bound = {n for n, _, _ in tng.EXPORTS}
    assert bound <= set(names)


def _host_mesh(L, Lr, rmin, rmax, ls, rs):
    lib = tng.load_library()
    h = ctypes.c_void_p()
    st = lib.tng_mesh_create_uniform(None, L, Lr, rmin, rmax, ls, rs, ctypes.byref(h))
    assert st == 0, lib.tng_last_error(None)
    info = tng._MeshInfo()
    assert lib.tng_mesh_get_info(h, ctypes.byref(info)) == 0
    shape = (info.num_subdomains, info.nodes_lateral, info.nodes_lateral, info.nodes_radial)

    def cp(fn, sh, dt):
        out = np.empty(sh, dtype=dt)
        assert getattr(lib, fn)(h, out.ctypes.data_as(ctypes.c_void_p)) == 0
        return out

    res = dict(info=info, shape=shape,
               coords=cp("tng_mesh_copy_coords", shape[:3] + (3,), np.float64),
               radii=cp("tng_mesh_copy_radii", (shape[0], shape[3]), np.float64),
               bmask=cp("tng_mesh_copy_boundary_mask", shape, np.uint8),
               omask=cp("tng_mesh_copy_ownership_mask", shape, np.uint8))
    lib.tng_mesh_destroy(h)
    return res


@pytest.mark.parametrize("L,Lr,ls,rs", [(1, 1, 0, 0), (2, 3, 0, 0), (4, 4, 0, 0), (3, 3, 1, 1), (4, 3, 2, 1),
                                        (6, 6, 0, 0)])
def test_host_mesh_bit_exact_vs_oracle(L, Lr, ls, rs):
    h = _host_mesh(L, Lr, 0.55, 1.0, ls, rs)
    m = O.Mesh(L, Lr, 0.55, 1.0, ls, rs)
    assert h["shape"] == m.shape
    assert np.array_equal(h["coords"].view(np.uint64), m.coords.view(np.uint64))
    assert np.array_equal(h["radii"].view(np.uint64), m.radii.view(np.uint64))
    assert np.array_equal(h["bmask"], m.boundary_mask)
    assert np.array_equal(h["omask"], m.ownership_mask)
    assert h["info"].num_owned == (10 * 4 ** L + 2) * (2 ** Lr + 1)


def test_invalid_arguments_are_reported():
    lib = tng.load_library()
    h = ctypes.c_void_p()
    assert lib.tng_mesh_create_uniform(None, 0, 2, 0.5, 1.0, 0, 0, ctypes.byref(h)) == 1
    assert b"level" in lib.tng_last_error(None)
    assert lib.tng_mesh_create_uniform(None, 2, 2, 1.0, 0.5, 0, 0, ctypes.byref(h)) == 1
    assert lib.tng_op_apply(None, None, None) == 1
