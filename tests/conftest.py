import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "terra-ng_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def consistent(mesh, field):
    """Copy the OWNED value of every node to all of its copies (oracle mesh)."""
    cid = mesh.canonical_ids.ravel()
    own = mesh.ownership_mask.ravel() == 1
    f = np.asarray(field, dtype=np.float64)
    lead = f.shape[: f.ndim - 4]
    flat = f.reshape((-1, mesh.nalloc))
    val = np.zeros((flat.shape[0], mesh.nunique))
    val[:, cid[own]] = flat[:, own]
    return np.ascontiguousarray(val[:, cid].reshape(lead + mesh.shape))


def rel_err(a, b, mask=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if mask is not None:
        a = a[..., mask]
        b = b[..., mask]
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="session")
def gpu_ctx():
    import tng

    return tng.default_context()
