"""GPU parity of the Robin / port surface assembly (robin.hip, emerge_amd.robin) against the
oracle restatement of robinbc.py, and of the CSC add (K += generate_csc(B)) against SciPy."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from oracle import robin as orob
from tests.conftest import ref_tensor, rel_fro

pytestmark = pytest.mark.gpu

C0 = 299792458.0
A_WG, B_WG, L_WG = 22.86e-3, 10.16e-3, 50e-3


def _port_tris(field, z):
    m = field.mesh
    bt = m.boundary_tris()
    zc = m.nodes[2][m.tris[:, bt]]
    return bt[np.all(np.abs(zc - z) < 1e-9, axis=0)]


def _port_mode(z):
    from emerge_amd.ports import LocalFrame, RectangularWaveguideMode
    fr = LocalFrame(origin=[A_WG / 2, B_WG / 2, z], xax=[1, 0, 0], yax=[0, 1, 0], zax=[0, 0, 1])
    return RectangularWaveguideMode(fr, A_WG, B_WG, (1, 0))


@pytest.mark.parametrize("which", ["wr90", "delaunay"])
def test_blocks_match_oracle(which, wr90_field, delaunay_field):
    from emerge_amd.assembly import Ned2Plan
    from emerge_amd.robin import RobinSurface
    f = wr90_field if which == "wr90" else delaunay_field
    surf = f.mesh.boundary_tris()
    plan = Ned2Plan.from_field(f)
    rs = RobinSurface(plan, f, surf)
    g = 0.7 + 158.2j
    by_face = rs.blocks(g, by_face=True).cpu().numpy()
    ref = orob.bc_entries(f.mesh.nodes, f.mesh.tris, surf, g)
    assert rel_fro(by_face, ref) <= 1e-12
    comp = rs.blocks(g).cpu().numpy().reshape(-1, 64)
    assert np.array_equal(comp, by_face.reshape(-1, 64)[surf])
    # accumulation into an existing Bmat (the reference adds, robinbc.py:400)
    import torch
    base = torch.as_tensor(ref.copy(), device="cuda")
    rs.blocks(g, out=base, by_face=True)
    assert rel_fro(base.cpu().numpy(), 2 * ref) <= 1e-12


def test_dropin_robin_bc(wr90_field):
    from emerge_amd.robin import assemble_robin_bc
    f = wr90_field
    surf = _port_tris(f, 0.0)
    B0 = np.zeros(f.mesh.n_tris * 64, complex)
    B = assemble_robin_bc(f, B0, surf, 1j * 158.238)
    assert rel_fro(B, orob.bc_entries(f.mesh.nodes, f.mesh.tris, surf, 1j * 158.238)) <= 1e-12


@pytest.mark.parametrize("form", ["plain", "scattered"])
def test_force_matches_oracle(wr90_field, form):
    from emerge_amd.robin import assemble_robin_bc_bvec, assemble_robin_bc_bvec_scat
    from emerge_amd.robin import generate_points_3d
    f = wr90_field
    surf = _port_tris(f, 0.0)
    md = _port_mode(0.0)
    k0 = 2 * np.pi * 10e9 / C0
    Ufunc = lambda x, y, z: md.get_Uinc(x, y, z, k0)  # noqa: E731
    x, y, z = generate_points_3d(f.mesh.nodes, f.mesh.tris, surf)
    U = Ufunc(x, y, z).reshape(3, 6, len(surf))
    if form == "plain":
        b = assemble_robin_bc_bvec(f, surf, Ufunc)
        ref = orob.force_entries(f.mesh.nodes, f.mesh.tris, f.tri_to_field, f.n_field, surf, U)
    else:
        rng = np.random.default_rng(0)
        Cf = lambda x, y, z: (rng.normal(size=(3, x.size)) + 0j)  # noqa: E731
        C = Cf(x, y, z).reshape(3, 6, len(surf))
        normals = np.tile(np.array([[0.0], [0.0], [-1.0]]), (1, len(surf)))
        b = assemble_robin_bc_bvec_scat(f, surf, Ufunc, lambda x, y, z: C.reshape(3, -1),
                                        normals)
        ref = orob.force_entries(f.mesh.nodes, f.mesh.tris, f.tri_to_field, f.n_field, surf, U,
                                 U_curl=C, normals=normals)
    assert rel_fro(b, ref) <= 1e-12
    assert np.count_nonzero(ref) > 0


@pytest.mark.parametrize("restricted", [False, True])
def test_csc_add_equals_scipy(wr90_field, restricted):
    """K + generate_csc(gamma B) on the device equals SciPy's sum of the same matrices,
    for the full CSC and the PEC-restricted block."""
    import torch
    from emerge_amd.assembly import CscMap, Ned2Plan
    from emerge_amd.pec import pec_solve_ids
    from emerge_amd.robin import RobinSurface
    f = wr90_field
    T = f.n_tets
    k0 = 2 * np.pi * 10e9 / C0
    plan = Ned2Plan.from_field(f)
    cm = CscMap(plan)
    ports = np.concatenate([_port_tris(f, 0.0), _port_tris(f, L_WG)])
    rs = RobinSurface(plan, f, ports)
    unit = rs.blocks(1.0)
    gamma = _port_mode(0.0).get_gamma(k0)
    K, M = plan.assemble_km(1.0, 1.0, "fp64")
    # reference side: oracle CSC + scipy COO of the oracle's blocks (femdata.generate_csc)
    I = ref_tensor(T, "identity", None)
    dE, dB, r, c = oracle.tet_mass_stiffness_matrices(f, I, I, [])
    A_ref = oracle.CSCMap(r, c, f.n_field).to_csc(dE - k0 ** 2 * dB)
    Bm = orob.bc_entries(f.mesh.nodes, f.mesh.tris, ports, gamma)
    tri_r = np.repeat(f.tri_to_field.T[:, :, None], 8, axis=2).reshape(-1)
    tri_c = np.repeat(f.tri_to_field.T[:, None, :], 8, axis=1).reshape(-1)
    nz = np.flatnonzero(Bm != 0)
    R = sp.csc_matrix((Bm[nz], (tri_r[nz], tri_c[nz])), shape=A_ref.shape)
    full_ref = (A_ref + R).tocsc()
    if not restricted:
        data = cm.to_csc_combine(K, M, k0)
        pos = rs.positions(cm)
        assert int((pos < 0).sum()) == 0
        rs.add_to_csc(data, pos, unit, gamma)
        ex = cm.export("indptr", "indices")
        got = sp.csc_matrix((data.cpu().numpy(), ex["indices"].cpu().numpy(),
                             ex["indptr"].cpu().numpy()), shape=A_ref.shape)
        ref = full_ref
    else:
        walls = np.setdiff1d(f.mesh.boundary_tris(), ports)
        sel = pec_solve_ids(plan, f, [], walls)
        sub = cm.restrict(sel)
        data = sub.to_csc_combine(K, M, k0)
        pos = rs.positions(sub, sel)
        rs.add_to_csc(data, pos, unit, gamma)
        ex = sub.export("indptr", "indices")
        got = sp.csc_matrix((data.cpu().numpy(), ex["indices"].cpu().numpy(),
                             ex["indptr"].cpu().numpy()), shape=(sub.N, sub.N))
        ids = sel.solve_ids.cpu().numpy()
        ref = full_ref[:, ids][ids, :].tocsc()
    d = (got - ref)
    assert sp.linalg.norm(d) <= 1e-12 * sp.linalg.norm(ref)


def test_tri_rowcol_matches_femdata_order(wr90_field):
    from emerge_amd.assembly import Ned2Plan
    from emerge_amd.robin import RobinSurface
    f = wr90_field
    surf = _port_tris(f, L_WG)
    rs = RobinSurface(Ned2Plan.from_field(f), f, surf)
    rows, cols = (t.cpu().numpy().reshape(-1, 8, 8) for t in rs.rowcol())
    t2f = f.tri_to_field[:, surf].T  # (n, 8)
    assert np.array_equal(rows, np.repeat(t2f[:, :, None], 8, axis=2))
    assert np.array_equal(cols, np.repeat(t2f[:, None, :], 8, axis=1))
