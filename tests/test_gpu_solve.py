"""End to end on the waveguide box (22.86 x 10.16 x 50 mm, 7,200 tets, f = 10 GHz): the
GPU-assembled and the oracle-assembled systems, after PEC elimination (the solver
slices A[solve_ids][:, solve_ids], solver.py:2071), give the same field for the same
excitation through SciPy's sparse direct solver; and the GPU matrices reproduce the
cavity's analytic eigenfrequencies."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

import oracle
from tests.conftest import ref_tensor, rel_fro
from tests.test_oracle import _pec_interior_dofs

pytestmark = pytest.mark.gpu

C0 = 299792458.0


def test_fields_agree_after_pec_elimination(wr90_field):
    from emerge_amd.curlcurl import tet_mass_stiffness_matrices
    f = wr90_field
    T = f.n_tets
    I = ref_tensor(T, "identity", None)
    k0 = 2 * np.pi * 10e9 / C0
    Evec, Bvec, cscmap = tet_mass_stiffness_matrices(f, I, I, np.array([]))
    A_gpu = cscmap.to_csc(Evec - Bvec * k0 ** 2)
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(f, I, I, [])
    A_ref = oracle.CSCMap(rows, cols, f.n_field).to_csc(dE - dB * k0 ** 2)
    keep = _pec_interior_dofs(f)
    rng = np.random.default_rng(0)
    b = np.zeros(f.n_field, complex)
    # a localised excitation: a few code-64 DOFs near the box centre
    b[rng.choice(keep[keep < f.mesh.n_edges], 12, replace=False)] = 1.0 - 0.5j
    x_gpu = sla.spsolve(A_gpu[keep][:, keep].tocsc(), b[keep])
    x_ref = sla.spsolve(A_ref[keep][:, keep].tocsc(), b[keep])
    assert rel_fro(x_gpu, x_ref) < 1e-9
    # the solve itself is accurate: residual of the GPU system
    r = A_gpu[keep][:, keep] @ x_gpu - b[keep]
    assert np.linalg.norm(r) < 1e-9 * np.linalg.norm(b)


def test_gpu_matrices_give_cavity_eigenfrequencies(wr90_field):
    import torch
    from emerge_amd.assembly import CscMap, Ned2Plan
    f = wr90_field
    plan = Ned2Plan.from_field(f)
    K, M = plan.assemble_km(1.0, 1.0, "fp64")
    cm = CscMap(plan)
    from emerge_amd.curlcurl import CSCMapping
    m = CSCMapping(cm)
    Kc = m.to_csc(K).real
    Mc = m.to_csc(M).real
    keep = _pec_interior_dofs(f)
    sigma = (2 * np.pi * 9e9 / C0) ** 2
    vals = sla.eigsh(Kc[keep][:, keep].tocsc(), k=3, M=Mc[keep][:, keep].tocsc(), sigma=sigma,
                     which="LM", return_eigenvectors=False)
    fr = np.sort(np.sqrt(vals) * C0 / (2 * np.pi)) / 1e9
    assert np.all(np.abs(fr - np.array([7.20997, 8.88517, 11.13032])) / fr < 2e-3), fr
