"""GPU parity of the PEC elimination and solver hand-off (emerge_amd.pec, dof_select.hip)
against the oracle restatement of assembler.py:533-564/742 and SciPy's slicing
A[:, solve_ids][solve_ids, :] (solver.py:2071)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle
from oracle import pec as opec
from tests.conftest import ref_tensor, rel_fro

pytestmark = pytest.mark.gpu

C0 = 299792458.0


def _oracle_ids(f, cond, tris):
    s = opec.pec_dof_set(f.tet_to_field, cond, tris, f.mesh.tri_to_edge, f.edge_to_field,
                         f.tri_to_field)
    return opec.solve_ids(f.n_field, s)


@pytest.mark.parametrize("case", ["walls", "conductors", "both", "none"])
def test_pec_solve_ids_match_oracle(wr90_field, delaunay_field, case):
    from emerge_amd.assembly import Ned2Plan
    from emerge_amd.pec import pec_solve_ids
    f = wr90_field if case in ("walls", "none") else delaunay_field
    rng = np.random.default_rng(7)
    cond = [] if case in ("walls", "none") else rng.choice(f.n_tets, 40, replace=False)
    tris = [] if case in ("conductors", "none") else f.mesh.boundary_tris()
    if case == "both":
        tris = rng.choice(f.mesh.boundary_tris(), 300, replace=False)
    plan = Ned2Plan.from_field(f)
    sel = pec_solve_ids(plan, f, cond, tris)
    ref = _oracle_ids(f, cond, tris)
    assert np.array_equal(sel.solve_ids.cpu().numpy(), ref)
    ni = sel.new_index.cpu().numpy()
    exp = np.full(f.n_field, -1)
    exp[ref] = np.arange(len(ref))
    assert np.array_equal(ni, exp)


def test_pec_mask_rejects_bad_ids(wr90_field):
    from emerge_amd._lib import EmError
    from emerge_amd.assembly import Ned2Plan
    from emerge_amd.pec import pec_solve_ids
    f = wr90_field
    plan = Ned2Plan.from_field(f)
    with pytest.raises(EmError, match="conductor tet"):
        pec_solve_ids(plan, f, [f.n_tets], [])
    with pytest.raises(EmError, match="PEC face"):
        pec_solve_ids(plan, f, [], [-1])


def _device_csc(A, dev):
    import torch
    A = A.tocsc()
    A.sort_indices()
    return (torch.as_tensor(A.indptr.astype(np.int64), device=dev),
            torch.as_tensor(A.indices.astype(np.int64), device=dev),
            torch.as_tensor(A.data.astype(np.complex128), device=dev))


@pytest.mark.parametrize("drop_zeros", [False, True])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_select_csc_matches_scipy(drop_zeros, dtype):
    import torch
    from emerge_amd.pec import SolveIds, select_csc
    rng = np.random.default_rng(3)
    n = 5000
    A = sp.random(n, n, density=0.004, random_state=3, format="lil", dtype=np.complex128,
                  data_rvs=lambda k: rng.normal(size=k) + 1j * rng.normal(size=k))
    A[:, 11] = 0
    for c in range(0, n, 97):  # a few long columns: more than one wavefront pass
        A[rng.choice(n, 150, replace=False), c] = 1.0 + 2.0j
    A = A.tocsc()
    A.sort_indices()
    A.data[::13] = 0
    keep = rng.random(n) < 0.6
    ip, ix, dx = _device_csc(A, "cuda")
    if dtype == "c64":
        dx = dx.to(torch.complex64)
    sel = SolveIds.from_keep(torch.as_tensor(keep.astype(np.uint8), device="cuda"))
    ids = np.flatnonzero(keep)
    assert np.array_equal(sel.solve_ids.cpu().numpy(), ids)
    sp_, si, sd = select_csc(ip, ix, dx, sel, drop_zeros=drop_zeros)
    ref = A[:, ids][ids, :].tocsc()
    if drop_zeros:
        ref.eliminate_zeros()
    ref.sort_indices()
    assert np.array_equal(sp_.cpu().numpy(), ref.indptr)
    assert np.array_equal(si.cpu().numpy(), ref.indices)
    want = ref.data.astype(np.complex64 if dtype == "c64" else np.complex128)
    assert np.array_equal(sd.cpu().numpy(), want)
    # pattern only
    sp2, si2, sd2 = select_csc(ip, ix, None, sel)
    if not drop_zeros:
        assert sd2 is None and torch.equal(sp2, sp_) and torch.equal(si2, si)


def test_gather_scatter_rows():
    import torch
    from emerge_amd.pec import SolveIds, gather_rows, scatter_solution
    rng = np.random.default_rng(4)
    n = 10007
    keep = rng.random(n) < 0.8
    sel = SolveIds.from_keep(torch.as_tensor(keep, device="cuda"))
    ids = np.flatnonzero(keep)
    b = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    bs = gather_rows(torch.as_tensor(b, device="cuda"), sel).cpu().numpy()
    assert np.array_equal(bs, b[ids, :])
    x = rng.normal(size=len(ids)) + 1j * rng.normal(size=len(ids))
    full = scatter_solution(torch.as_tensor(x.astype(np.complex64), device="cuda"), sel)
    ref = np.zeros(n, np.complex64)
    ref[ids] = x.astype(np.complex64)
    assert np.array_equal(full.cpu().numpy(), ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_restricted_map_equals_selected_full(delaunay_field, precision):
    """The restricted map's to_csc_combine gives the free-DOF block of the full map's
    result bit for bit (same contributors, same summation order), and matches the
    oracle's CSC sliced by SciPy."""
    import torch
    from emerge_amd.assembly import CscMap, Ned2Plan
    from emerge_amd.pec import pec_solve_ids, select_csc
    f = delaunay_field
    rng = np.random.default_rng(5)
    T = f.n_tets
    cond = rng.choice(T, 100, replace=False)
    tris = f.mesh.boundary_tris()
    plan = Ned2Plan.from_field(f)
    eps = torch.tensor(rng.uniform(1, 10, T) * (1 - 0.01j), dtype=torch.complex128)
    mu = torch.tensor(rng.uniform(1, 2, (T, 3)), dtype=torch.complex128)
    K, M = plan.assemble_km(eps, mu, precision)
    cm = CscMap(plan)
    sel = pec_solve_ids(plan, f, cond, tris)
    sub = cm.restrict(sel)
    assert sub.N == sel.n_solve
    k0 = 2 * np.pi * 9e9 / C0
    full = cm.to_csc_combine(K, M, k0)
    ex = cm.export("indptr", "indices")
    sp_, si, sd = select_csc(ex["indptr"], ex["indices"], full, sel)
    blk = sub.to_csc_combine(K, M, k0)
    exs = sub.export("indptr", "indices")
    assert torch.equal(exs["indptr"], sp_) and torch.equal(exs["indices"], si)
    assert torch.equal(blk, sd)
    # the restricted map's incidence gather = its per-entry segment gather, plain and mirrored
    assert torch.equal(blk, sub.to_csc_combine(K, M, k0, segment_gather=True))
    assert torch.equal(sub.to_csc_combine(K, M, k0, block_symmetric=True),
                       sub.to_csc_combine(K, M, k0, block_symmetric=True, segment_gather=True))
    one = sub.to_csc(K)
    assert torch.equal(one, select_csc(ex["indptr"], ex["indices"], cm.to_csc(K), sel)[2])
    # against the oracle, assembled and sliced on the CPU
    er = np.zeros((3, 3, T), complex)
    ur = np.zeros((3, 3, T), complex)
    for d in range(3):
        er[d, d] = eps.numpy()
        ur[d, d] = mu.numpy()[:, d]
    dE, dB, r, c = oracle.tet_mass_stiffness_matrices(f, er, ur, [], local_frame=True)
    A_ref = oracle.CSCMap(r, c, f.n_field).to_csc(dE - k0 ** 2 * dB)
    ids = _oracle_ids(f, cond, tris)
    ref = A_ref[:, ids][ids, :].tocsc()
    ref.sort_indices()
    assert np.array_equal(exs["indptr"].cpu().numpy(), ref.indptr)
    assert np.array_equal(exs["indices"].cpu().numpy(), ref.indices)
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert rel_fro(blk.cpu().numpy().astype(complex), ref.data) <= tol
    with pytest.raises(Exception):
        sub.export("perm")


def test_cavity_eigenfrequencies_through_gpu_elimination(wr90_field):
    """PEC walls collected and eliminated on the GPU (restricted maps of K and M):
    TE101/102/103 of the 22.86 x 10.16 x 50 mm cavity."""
    from emerge_amd.assembly import CscMap, Ned2Plan
    from emerge_amd.pec import pec_solve_ids, to_scipy
    f = wr90_field
    plan = Ned2Plan.from_field(f)
    K, M = plan.assemble_km(1.0, 1.0, "fp64")
    sel = pec_solve_ids(plan, f, [], f.mesh.boundary_tris())
    sub = CscMap(plan).restrict(sel)
    ex = sub.export("indptr", "indices")
    Ks = to_scipy(ex["indptr"], ex["indices"], sub.to_csc(K), sub.N).real
    Ms = to_scipy(ex["indptr"], ex["indices"], sub.to_csc(M), sub.N).real
    sigma = (2 * np.pi * 9e9 / C0) ** 2
    vals = sla.eigsh(Ks, k=3, M=Ms, sigma=sigma, which="LM", return_eigenvectors=False)
    fr = np.sort(np.sqrt(vals) * C0 / (2 * np.pi)) / 1e9
    assert np.all(np.abs(fr - np.array([7.20997, 8.88517, 11.13032])) / fr < 2e-3), fr


def test_select_system_solve_and_scatter(wr90_field):
    """select_system + scatter_solution around a SciPy solve reproduce the host-side
    slicing path: same field on every DOF, zeros on the PEC DOFs."""
    import torch
    from emerge_amd.assembly import CscMap, Ned2Plan
    from emerge_amd.pec import pec_solve_ids, scatter_solution, select_system, to_scipy
    f = wr90_field
    plan = Ned2Plan.from_field(f)
    k0 = 2 * np.pi * 10e9 / C0
    A = plan.assemble_a(1.0, 1.0, k0, "fp64")
    cm = CscMap(plan)
    ex = cm.export("indptr", "indices")
    Ad = (ex["indptr"], ex["indices"], cm.to_csc(A))
    sel = pec_solve_ids(plan, f, [], f.mesh.boundary_tris())
    ids = sel.solve_ids.cpu().numpy()
    rng = np.random.default_rng(0)
    b = np.zeros(f.n_field, complex)
    b[rng.choice(ids[ids < f.n_edges], 12, replace=False)] = 1.0 - 0.5j
    (ip, ix, dx), bsel = select_system(Ad, b, sel)
    As = to_scipy(ip, ix, dx, sel.n_solve)
    x = sla.spsolve(As, bsel.cpu().numpy())
    sol = scatter_solution(torch.as_tensor(x, device="cuda"), sel).cpu().numpy()
    Ah = to_scipy(*Ad, f.n_field)
    x_ref = sla.spsolve(Ah[:, ids][ids, :].tocsc(), b[ids])
    ref = np.zeros(f.n_field, complex)
    ref[ids] = x_ref
    assert rel_fro(sol, ref) < 1e-12
