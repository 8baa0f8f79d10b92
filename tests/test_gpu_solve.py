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


def test_delaunay_lossy_materials_field_error_mixed_vs_fp64():
    """Per-solved-field error of the mixed path with the Delaunay configuration's materials
    (per-tet complex scalar eps_r, Re in [1, 10], loss tangent in [0, 0.02]; real diagonal
    mu_r in [1, 2]) at 10 GHz. The mesh is a seeded 300-point Delaunay box scaled to
    1.5 cm (3,029 tets, about five second-order elements per wavelength at eps_r = 10),
    with PEC walls (GPU elimination) and a seeded random excitation. The FP64 and mixed
    GPU systems are solved by the same SciPy direct solver."""
    import torch
    from emerge_amd.assembly import CscMap, Ned2Plan
    from emerge_amd.mesh import Nedelec2Field, TetMesh, delaunay_box_mesh
    from emerge_amd.pec import pec_solve_ids, to_scipy
    m0 = delaunay_box_mesh(300, seed=12)
    f = Nedelec2Field(TetMesh(m0.nodes * 0.015, m0.tets))
    T = f.n_tets
    rng = np.random.default_rng(3)
    er, ur = ref_tensor(T, "scalar", rng), ref_tensor(T, "diag", rng, complex_=False)
    k0 = 2 * np.pi * 10e9 / C0
    plan = Ned2Plan.from_field(f)
    sel = pec_solve_ids(plan, f, [], f.mesh.boundary_tris())
    sub = CscMap(plan).restrict(sel)
    ex = sub.export("indptr", "indices")
    K, M = plan.assemble_km(er, ur, "fp64")
    a64 = sub.to_csc_combine(K, M, k0, block_symmetric=True)
    Km, Mm = plan.assemble_km(er, ur, "mixed")
    a32 = sub.to_csc_combine(Km, Mm, k0, block_symmetric=True).to(torch.complex128)
    n = sel.n_solve
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    x64 = sla.spsolve(to_scipy(ex["indptr"], ex["indices"], a64, n), b)
    x32 = sla.spsolve(to_scipy(ex["indptr"], ex["indices"], a32, n), b)
    e_mat = rel_fro(a32.cpu().numpy(), a64.cpu().numpy())
    e_fld = rel_fro(x32, x64)
    print(f"\n[Delaunay 3,029 tets, lossy eps_r, diagonal mu_r, 10 GHz, {n} free DOFs] mixed vs "
          f"FP64: matrix {e_mat:.3e}, solved field {e_fld:.3e}")
    assert e_mat < 1e-6 and e_fld < 1e-3
