"""GPU parity of the heat-conduction (Lagrange-2) assembly (heat.hip, emerge_amd.heat)
against the oracle restatement of grad.py, the CSC pattern against SciPy, and a global
energy pin of the assembled conductance matrix."""
import time

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import heat as oh
from tests.conftest import rel_fro

pytestmark = pytest.mark.gpu


def _field(which, wr90_field, delaunay_field):
    from emerge_amd.mesh import Legrange2Field
    return Legrange2Field((wr90_field if which == "wr90" else delaunay_field).mesh)


def _cond(T, rng, kind):
    C = np.zeros((3, 3, T))
    if kind == "scalar":
        v = rng.uniform(1, 5, T)
        for d in range(3):
            C[d, d] = v
    elif kind == "diag":
        for d in range(3):
            C[d, d] = rng.uniform(1, 5, T)
    else:
        A = rng.normal(size=(3, 3, T)) * 0.3
        C = A + np.eye(3)[:, :, None] * 3
    return C


@pytest.mark.parametrize("which", ["wr90", "delaunay"])
@pytest.mark.parametrize("kind", ["scalar", "diag", "full"])
def test_heat_blocks_match_oracle(which, kind, wr90_field, delaunay_field):
    import torch
    from emerge_amd.heat import HeatPlan
    f = _field(which, wr90_field, delaunay_field)
    m = f.mesh
    T = f.n_tets
    rng = np.random.default_rng(5)
    C = _cond(T, rng, kind)
    rho = rng.uniform(1e6, 4e6, T)
    plan = HeatPlan.from_field(f)
    if kind == "scalar":
        cond = torch.as_tensor(C[0, 0].copy())
    elif kind == "diag":
        cond = torch.as_tensor(np.stack([C[d, d] for d in range(3)], axis=1).copy())
    else:
        cond = torch.as_tensor(C)  # reference (3, 3, T) layout
    K, M = plan.assemble(cond=cond, rho_cp=torch.as_tensor(rho))
    Kr, Mr, rows, cols = oh.build(m.nodes, m.tets, f.tet_to_field, cond=C, rho_cp=rho,
                                  rowcol=True, local_frame=(which == "delaunay"))
    assert rel_fro(K.cpu().numpy(), Kr) <= 1e-12
    assert rel_fro(M.cpu().numpy(), Mr) <= 1e-12
    r, c = plan.rowcol()
    assert np.array_equal(r.cpu().numpy(), rows) and np.array_equal(c.cpu().numpy(), cols)


@pytest.mark.parametrize("every", [3, 1])
def test_negative_orientation_tets(delaunay_field, every):
    """Tets stored with negative orientation follow the reference's |V| normalisation:
    every third tet flipped (chunks of mixed orientation: threads switch coefficient sets
    between chunks) and every tet flipped (every chunk negatively oriented)."""
    import torch
    from emerge_amd.heat import HeatPlan
    from emerge_amd.mesh import Legrange2Field, TetMesh
    m0 = delaunay_field.mesh
    tets = m0.tets.copy()
    flip = np.arange(0, m0.n_tets, every)
    tets[2, flip], tets[3, flip] = m0.tets[3, flip], m0.tets[2, flip]
    f = Legrange2Field(TetMesh(m0.nodes, tets))
    assert np.sum(f.mesh.signed_dets() < 0) == len(flip)
    T = f.n_tets
    rng = np.random.default_rng(6)
    C = _cond(T, rng, "full")
    rho = rng.uniform(1, 2, T)
    K, M = HeatPlan.from_field(f).assemble(cond=torch.as_tensor(C), rho_cp=torch.as_tensor(rho))
    Kr, Mr, _, _ = oh.build(f.mesh.nodes, f.mesh.tets, f.tet_to_field, cond=C, rho_cp=rho,
                            local_frame=True)
    assert rel_fro(K.cpu().numpy(), Kr) <= 1e-12
    assert rel_fro(M.cpu().numpy(), Mr) <= 1e-12


def test_heat_csc_and_energy(wr90_field):
    """CSC of the conductance equals SciPy's COO->CSC; a linear temperature u = g . x
    (in the Lagrange-2 space) has energy u^T K u = c |g|^2 vol * sum(w) exactly, and K
    annihilates constants."""
    import torch
    from emerge_amd.heat import HeatCscMap, HeatPlan, tet_mass_matrix, tet_stiffness_matrix
    from emerge_amd.mesh import Legrange2Field
    f = Legrange2Field(wr90_field.mesh)
    m = f.mesh
    T = f.n_tets
    c = 2.5
    plan = HeatPlan.from_field(f)
    K, M = plan.assemble(cond=torch.tensor(c, dtype=torch.float64),
                         rho_cp=torch.ones(T, dtype=torch.float64))
    cm = HeatCscMap(plan)
    r, cc = plan.rowcol()
    S = sp.coo_matrix((K.cpu().numpy(), (r.cpu().numpy(), cc.cpu().numpy())),
                      shape=(f.n_field, f.n_field)).tocsc()
    S.sum_duplicates()
    S.sort_indices()
    ex = cm.export("indptr", "indices")
    assert np.array_equal(ex["indptr"].cpu().numpy(), S.indptr)
    assert np.array_equal(ex["indices"].cpu().numpy(), S.indices)
    vals = cm.to_csc(K).cpu().numpy()
    assert rel_fro(vals, S.data) <= 1e-14
    A = sp.csc_matrix((vals, S.indices, S.indptr), shape=S.shape)
    g = np.array([0.3, -1.2, 0.7])
    xn = m.nodes
    u = np.concatenate([g @ xn, g @ (0.5 * (xn[:, m.edges[0]] + xn[:, m.edges[1]]))])
    vol = 22.86e-3 * 10.16e-3 * 50e-3
    wsum = oh.REFERENCE_RULE[0].sum()
    assert abs(u @ (A @ u) - c * (g @ g) * vol * wsum) <= 1e-10 * c * (g @ g) * vol
    assert np.abs(A @ np.ones(f.n_field)).max() <= 1e-10 * np.abs(vals).max()
    # transient combine K + M / dt in one pass
    dt = 1e-3
    both = cm.to_csc(K, M, 1 / dt).cpu().numpy()
    assert rel_fro(both, vals + cm.to_csc(M).cpu().numpy() / dt) <= 1e-14
    # drop-in entry points (grad.py:417-462)
    Kd, mapping = tet_stiffness_matrix(f, np.full((3, 3, T), 0.0) + c * np.eye(3)[:, :, None])
    assert rel_fro(Kd, K.cpu().numpy()) <= 1e-14
    assert (mapping.to_csc(Kd) - A).nnz == 0 or sp.linalg.norm(mapping.to_csc(Kd) - A) <= 1e-12
    Md, _ = tet_mass_matrix(f, np.ones(T))
    assert rel_fro(Md, M.cpu().numpy()) <= 1e-14


@pytest.mark.parametrize("n", [205, 1228, 1229, 2000, 6553])
def test_heat_long_columns_and_the_limit(n):
    """The Lagrange-2 DOFs of a fan's axis (its two vertices and its midpoint) lie in all
    n tets: columns of 10 n COO entries, past the column kernel's 2,048-entry LDS sort
    from 205 tets on (the long-column kernel builds them; rows in LDS up to 1,228 tets,
    in global scratch beyond), up to 6,553 tets. Pattern and values equal SciPy's
    COO->CSC with duplicates summed, the incidence gather equals the segment gather, and
    6,554 tets are refused with EM_EUNSUPPORTED."""
    import torch
    from emerge_amd._lib import UnsupportedError
    from emerge_amd.heat import HeatCscMap, HeatPlan
    from emerge_amd.mesh import Legrange2Field
    from tests.test_gpu_csc import _fan_field
    f = Legrange2Field(_fan_field(n).mesh)
    T = f.n_tets
    plan = HeatPlan.from_field(f)
    rng = np.random.default_rng(n)
    K, M = plan.assemble(cond=torch.tensor(rng.uniform(1, 3, T)),
                         rho_cp=torch.tensor(rng.uniform(1, 2, T)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cm = HeatCscMap(plan)
    torch.cuda.synchronize()
    print(f"heat map of a {n}-tet fan ({10 * n} entries in the axis columns): "
          f"{(time.perf_counter() - t0) * 1e3:.1f} ms", flush=True)
    r, cc = plan.rowcol()
    S = sp.coo_matrix((K.cpu().numpy(), (r.cpu().numpy(), cc.cpu().numpy())),
                      shape=(f.n_field, f.n_field)).tocsc()
    S.sum_duplicates()
    S.sort_indices()
    ex = cm.export("indptr", "indices")
    assert np.array_equal(ex["indptr"].cpu().numpy(), S.indptr)
    assert np.array_equal(ex["indices"].cpu().numpy(), S.indices)
    vals = cm.to_csc(K)
    assert rel_fro(vals.cpu().numpy(), S.data) <= 1e-14
    assert torch.equal(vals, cm.to_csc(K, segment_gather=True))
    assert torch.equal(cm.to_csc(K, M, 1e3), cm.to_csc(K, M, 1e3, segment_gather=True))
    if n == 2000:
        bigger = HeatPlan.from_field(Legrange2Field(_fan_field(6554).mesh))
        with pytest.raises(UnsupportedError, match="6554 tets \\(at most 6553"):
            HeatCscMap(bigger)


def test_heat_csc_incidence_gather_equals_segment_gather(delaunay_field):
    """The column-incidence gather of real data sums every CSC entry's contributors in COO
    order, like the per-entry segment gather: the results are equal bit for bit, with and
    without the fused K + s M combine, on random (not symmetric) block values."""
    import torch
    from emerge_amd.heat import HeatCscMap, HeatPlan
    from emerge_amd.mesh import Legrange2Field
    f = Legrange2Field(delaunay_field.mesh)
    plan = HeatPlan.from_field(f)
    cm = HeatCscMap(plan)
    g = torch.Generator().manual_seed(11)
    E = torch.randn(cm.nnz_coo, generator=g, dtype=torch.float64).cuda()
    B = torch.randn(cm.nnz_coo, generator=g, dtype=torch.float64).cuda()
    for args in ((E,), (E, B, 1e3)):
        inc = cm.to_csc(*args)
        seg = cm.to_csc(*args, segment_gather=True)
        assert torch.equal(inc, seg)
    # and against the sum in sorted-contributor order on the host
    ex = cm.export("perm", "target_csc")
    perm, tgt = ex["perm"].cpu().numpy(), ex["target_csc"].cpu().numpy()
    q = tgt  # CSC entry of the k-th contributor in sorted order (csc_cast.py:162-177)
    assert np.all(np.diff(q) >= 0)
    segs = np.searchsorted(q, np.arange(cm.nnz + 1))
    vals = E.cpu().numpy()[perm]
    acc = np.zeros(len(segs) - 1)
    lens = np.diff(segs)
    for r in range(int(lens.max())):
        m = lens > r
        acc[m] += vals[segs[:-1][m] + r]
    assert np.array_equal(acc, cm.to_csc(E).cpu().numpy())


def test_heat_blocks_bit_symmetric_and_mirrored_gather(delaunay_field):
    """With a diagonal conductivity every conductance block is bit-symmetric (the 10-term
    form adds the off-diagonal coefficient pairs in one order for (i, j) and (j, i)), as is
    every capacity block; the mirrored incidence gather then equals the segment gather bit
    for bit."""
    import torch
    from emerge_amd.heat import HeatCscMap, HeatPlan
    from emerge_amd.mesh import Legrange2Field
    f = Legrange2Field(delaunay_field.mesh)
    T = f.n_tets
    rng = np.random.default_rng(8)
    plan = HeatPlan.from_field(f)
    K, M = plan.assemble(cond=torch.as_tensor(rng.uniform(1, 5, (T, 3))),
                         rho_cp=torch.as_tensor(rng.uniform(1, 2, T)))
    for X in (K, M):
        b = X.reshape(T, 10, 10)
        assert torch.equal(b, b.transpose(1, 2))
    cm = HeatCscMap(plan)
    ref = cm.to_csc(K, M, 250.0, segment_gather=True)
    assert torch.equal(cm.to_csc(K, M, 250.0, block_symmetric=True), ref)
    assert torch.equal(cm.to_csc(K, block_symmetric=True), cm.to_csc(K, segment_gather=True))


@pytest.mark.parametrize("kind", ["diag", "full"])
def test_heat_large_mesh_32_chunk_workgroups(kind):
    """A mesh above 65,536 output chunks (~390k tets) runs 32 chunks per workgroup: a
    strided sample of tets (every 37th, and the last) matches the oracle, with every fifth
    tet negatively oriented so that threads switch coefficient sets between chunks, and M
    alone (no conductivity) equals M of the joint call bit for bit."""
    import torch
    from emerge_amd.heat import HeatPlan
    from emerge_amd.mesh import Legrange2Field, TetMesh, delaunay_mesh_with_tets
    m0 = delaunay_mesh_with_tets(390_000, seed=77)
    tets = m0.tets.copy()
    flip = np.arange(0, m0.n_tets, 5)
    tets[2, flip], tets[3, flip] = m0.tets[3, flip], m0.tets[2, flip]
    f = Legrange2Field(TetMesh(m0.nodes, tets))
    T = f.n_tets
    assert T * 100 >= 65536 * 512, T
    rng = np.random.default_rng(8)
    C = _cond(T, rng, kind)
    rho = rng.uniform(1, 2, T)
    plan = HeatPlan.from_field(f)
    K, M = plan.assemble(cond=torch.as_tensor(C), rho_cp=torch.as_tensor(rho))
    _, M_only = plan.assemble(rho_cp=torch.as_tensor(rho))
    assert torch.equal(M, M_only)
    ids = np.unique(np.concatenate([np.arange(0, T, 37), [T - 1]]))
    Kr, Mr, _, _ = oh.build(f.mesh.nodes, f.mesh.tets[:, ids], f.tet_to_field[:, ids],
                            cond=C[:, :, ids], rho_cp=rho[ids], local_frame=True)
    sel = (ids[:, None] * 100 + np.arange(100)[None, :]).ravel()
    sel_t = torch.as_tensor(sel, device=K.device)
    assert rel_fro(K[sel_t].cpu().numpy(), Kr) <= 1e-12
    assert rel_fro(M[sel_t].cpu().numpy(), Mr) <= 1e-12
