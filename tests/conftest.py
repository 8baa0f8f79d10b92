import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


HAS_GPU = _has_gpu()


def pytest_collection_modifyitems(config, items):
    if HAS_GPU:
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def wr90_field():
    """The waveguide parity mesh: 22.86 x 10.16 x 50 mm box, 10x5x24 hexes x 6 tets."""
    from emerge_amd.mesh import Nedelec2Field, hex_box_mesh
    return Nedelec2Field(hex_box_mesh(22.86e-3, 10.16e-3, 50e-3, 10, 5, 24))


@pytest.fixture(scope="session")
def delaunay_field():
    """Seeded unit-box Delaunay mesh, about 19k tets."""
    from emerge_amd.mesh import Nedelec2Field, delaunay_box_mesh
    return Nedelec2Field(delaunay_box_mesh(2500, seed=11))


def ref_tensor(T, kind, rng, complex_=True):
    """Random material tensors in the reference (3, 3, T) layout."""
    out = np.zeros((3, 3, T), np.complex128)
    if kind == "identity":
        for d in range(3):
            out[d, d] = 1.0
        return out
    if kind == "scalar":
        v = rng.uniform(1, 10, T) * (1 - 1j * rng.uniform(0, 0.02, T) * complex_)
        for d in range(3):
            out[d, d] = v
        return out
    if kind == "diag":
        for d in range(3):
            out[d, d] = rng.uniform(1, 2, T) + 1j * complex_ * rng.uniform(-0.1, 0.0, T)
        return out
    if kind == "sym":
        A = rng.normal(size=(3, 3, T)) * 0.2 + 1j * complex_ * rng.normal(size=(3, 3, T)) * 0.05
        out = A + A.transpose(1, 0, 2)
        for d in range(3):
            out[d, d] += 3.0
        return out
    if kind == "full":
        out = rng.normal(size=(3, 3, T)) * 0.3 + 1j * complex_ * rng.normal(size=(3, 3, T)) * 0.05
        for d in range(3):
            out[d, d] += 3.0
        return out
    raise ValueError(kind)


def rel_fro(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
