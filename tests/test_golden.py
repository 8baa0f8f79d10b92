"""The committed golden fixtures (exact integrals, tests/golden/make_fixtures.py) against
the oracle (CPU) and the HIP kernels (GPU)."""
import os

import numpy as np
import pytest

import oracle
from tests.conftest import rel_fro
from emerge_amd.mesh import Nedelec2Field, TetMesh

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _single():
    return np.load(os.path.join(HERE, "single_tet_24orders.npz"))


def _fields_24(g):
    for o, K, M in zip(g["orders"], g["K"], g["M"]):
        nodes = g["base"].T.copy()
        yield Nedelec2Field(TetMesh(nodes, o.reshape(4, 1))), K, M


def test_oracle_matches_single_tet_fixture():
    g = _single()
    er = g["eps"].reshape(3, 3, 1)
    ur = g["mu"].reshape(3, 3, 1)
    worst = 0.0
    for f, K, M in _fields_24(g):
        dE, dB, _, _ = oracle.tet_mass_stiffness_matrices(f, er, ur, np.array([]))
        worst = max(worst, rel_fro(dE.reshape(20, 20), K), rel_fro(dB.reshape(20, 20), M))
    assert worst < 1e-13


def test_oracle_matches_hex_fixture():
    g = np.load(os.path.join(HERE, "hex6.npz"))
    f = Nedelec2Field(TetMesh(g["nodes"], g["tets"]))
    I = np.zeros((3, 3, f.n_tets), complex)
    for d in range(3):
        I[d, d] = 1
    dE, dB, _, _ = oracle.tet_mass_stiffness_matrices(f, I, I, np.array([]))
    assert rel_fro(dE, g["K"].ravel()) < 1e-14 and rel_fro(dB, g["M"].ravel()) < 1e-14


@pytest.mark.gpu
def test_gpu_matches_single_tet_fixture():
    from emerge_amd.assembly import Ned2Plan
    g = _single()
    er = g["eps"].reshape(3, 3, 1)
    ur = g["mu"].reshape(3, 3, 1)
    worst = 0.0
    for f, K, M in _fields_24(g):
        Kg, Mg = Ned2Plan.from_field(f).assemble_km(er, ur, "fp64")
        worst = max(worst, rel_fro(Kg.cpu().numpy().reshape(20, 20), K),
                    rel_fro(Mg.cpu().numpy().reshape(20, 20), M))
    assert worst < 1e-13


@pytest.mark.gpu
def test_gpu_matches_hex_fixture():
    from emerge_amd.assembly import Ned2Plan
    g = np.load(os.path.join(HERE, "hex6.npz"))
    f = Nedelec2Field(TetMesh(g["nodes"], g["tets"]))
    Kg, Mg = Ned2Plan.from_field(f).assemble_km(1.0, 1.0, "fp64")
    assert rel_fro(Kg.cpu().numpy(), g["K"].ravel()) < 1e-14
    assert rel_fro(Mg.cpu().numpy(), g["M"].ravel()) < 1e-14
