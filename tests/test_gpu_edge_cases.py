"""GPU edge cases against the oracle.

* A mesh of one tet, in positive and negative orientation: Nédélec K and M, the fused
  A (FP64 and mixed), rows/cols, the CSC pattern and values (curlcurl.py:222-456,
  csc_cast.py:29-177), and the heat blocks and their CSC (grad.py:199-414). A single
  tet fills one partial tile and, for heat, one workgroup of a padded grid. The tet lies
  30 edge lengths from the origin, where the reference's absolute-coordinate cofactor
  sums leave ~1e-11 residues in entries that are zero; the oracle is evaluated in the
  tet's local frame (same algorithm, translated coordinates), as for Delaunay meshes.
* A plan whose every tet is a conductor (curlcurl.py:248-250 deletes them all): the
  reference returns empty COO arrays and an N x N CSC matrix without entries; so must
  every entry point here.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from oracle import heat as oh
from tests.conftest import ref_tensor, rel_fro

pytestmark = pytest.mark.gpu

K0 = 2 * np.pi * 10e9 / 299792458.0
# a 1 cm corner tet away from the origin
NODES = np.array([[0, 1.0, 0, 0], [0, 0, 1.0, 0], [0, 0, 0, 1.0]]) * 1e-2 + \
    np.array([[0.3], [0.1], [-0.2]])
ORDERS = [(0, 1, 2, 3), (1, 0, 2, 3)]  # det J > 0, det J < 0


def _one_tet(order):
    from emerge_amd.mesh import TetMesh
    return TetMesh(NODES, np.array(order, np.int64).reshape(4, 1))


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("kinds", [("scalar", "diag"), ("full", "full")])
def test_one_tet_nedelec(order, kinds):
    from emerge_amd.assembly import CscMap, Ned2Plan
    from emerge_amd.mesh import Nedelec2Field
    f = Nedelec2Field(_one_tet(order))
    rng = np.random.default_rng(11)
    er = ref_tensor(1, kinds[0], rng)
    ur = ref_tensor(1, kinds[1], rng)
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(f, er, ur, np.array([]),
                                                              local_frame=True)
    plan = Ned2Plan.from_field(f)
    K, M = plan.assemble_km(er, ur, "fp64")
    r, c = plan.rowcol()
    assert np.array_equal(r.cpu().numpy(), rows) and np.array_equal(c.cpu().numpy(), cols)
    assert rel_fro(K.cpu().numpy(), dE) <= 1e-12
    assert rel_fro(M.cpu().numpy(), dB) <= 1e-12
    Aref = dE - K0 ** 2 * dB
    assert rel_fro(plan.assemble_a(er, ur, K0, "fp64").cpu().numpy(), Aref) <= 1e-12
    Am = plan.assemble_a(er, ur, K0, "mixed").cpu().numpy().astype(np.complex128)
    assert rel_fro(Am, Aref) <= 1e-6
    cm = CscMap(plan)
    ref = oracle.CSCMap(rows, cols, f.n_field)
    ex = cm.export("indptr", "indices")
    assert np.array_equal(ex["indptr"].cpu().numpy(), ref.indptr)
    assert np.array_equal(ex["indices"].cpu().numpy(), ref.indices)
    vals = cm.to_csc_combine(K, M, K0).cpu().numpy()
    assert rel_fro(vals, ref.to_csc_data(Aref)) <= 1e-12
    # straight into CSC (no COO arrays): the one tet's block, in both orientations
    assert rel_fro(plan.assemble_csc(cm, er, ur, K0, "A", "fp64").cpu().numpy(),
                   ref.to_csc_data(Aref)) <= 1e-12
    assert rel_fro(plan.assemble_csc(cm, er, ur, 0.0, "M", "mixed").cpu().numpy()
                   .astype(np.complex128), ref.to_csc_data(dB)) <= 1e-6


@pytest.mark.parametrize("order", ORDERS)
def test_one_tet_heat(order):
    import torch
    from emerge_amd.heat import HeatCscMap, HeatPlan
    from emerge_amd.mesh import Legrange2Field
    m = _one_tet(order)
    f = Legrange2Field(m)
    rng = np.random.default_rng(12)
    A = rng.normal(size=(3, 3, 1)) * 0.3
    C = A + np.eye(3)[:, :, None] * 3  # full, non-symmetric conductivity
    rho = np.array([2.5e6])
    plan = HeatPlan.from_field(f)
    K, M = plan.assemble(cond=torch.as_tensor(C), rho_cp=torch.as_tensor(rho))
    Kr, Mr, rows, cols = oh.build(m.nodes, m.tets, f.tet_to_field, cond=C, rho_cp=rho,
                                  rowcol=True, local_frame=True)
    assert rel_fro(K.cpu().numpy(), Kr) <= 1e-12
    assert rel_fro(M.cpu().numpy(), Mr) <= 1e-12
    r, c = plan.rowcol()
    assert np.array_equal(r.cpu().numpy(), rows) and np.array_equal(c.cpu().numpy(), cols)
    dt = 1e-3
    S = sp.coo_matrix((Kr + Mr / dt, (rows, cols)), shape=(f.n_field, f.n_field)).tocsc()
    S.sum_duplicates()
    S.sort_indices()
    cm = HeatCscMap(plan)
    ex = cm.export("indptr", "indices")
    assert np.array_equal(ex["indptr"].cpu().numpy(), S.indptr)
    assert np.array_equal(ex["indices"].cpu().numpy(), S.indices)
    assert rel_fro(cm.to_csc(K, M, 1 / dt).cpu().numpy(), S.data) <= 1e-12


def test_plan_without_assembled_tets(wr90_field):
    import torch
    from emerge_amd.assembly import CscMap, Ned2Plan
    from emerge_amd.curlcurl import tet_mass_stiffness_matrices
    f = wr90_field
    T, N = f.n_tets, f.n_field
    rng = np.random.default_rng(13)
    er = ref_tensor(T, "scalar", rng)
    ur = ref_tensor(T, "identity", rng)
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(f, er, ur, np.arange(T))
    assert dE.size == 0 and rows.size == 0
    plan = Ned2Plan.from_field(f, tetids=np.zeros(0, np.int64))
    assert plan.n_assy == 0 and plan.nnz_coo == 0
    K, M = plan.assemble_km(er, ur, "fp64")
    assert K.numel() == 0 and M.numel() == 0
    assert plan.assemble_a(er, ur, K0, "mixed").numel() == 0
    assert tuple(plan.assemble_a(er, ur, np.array([K0, 2 * K0]), "fp64").shape) == (2, 0)
    plan.set_prefetch(4096)  # prefetched launches of nothing: no chunks, no launch
    assert plan.assemble_a(er, ur, K0, "mixed").numel() == 0
    assert all(x.numel() == 0 for x in plan.assemble_km(er, ur, "fp64"))
    r, c = plan.rowcol()
    assert r.numel() == 0 and c.numel() == 0
    cm = CscMap(plan)
    assert cm.nnz == 0
    ex = cm.export("indptr", "indices")
    assert np.array_equal(ex["indptr"].cpu().numpy(), np.zeros(N + 1, np.int64))
    assert cm.to_csc_combine(K, M, K0).numel() == 0
    assert cm.to_csc_combine(K, M, K0, block_symmetric=True).numel() == 0
    for which in ("A", "K", "M"):
        assert plan.assemble_csc(cm, er, ur, K0, which, "fp64").numel() == 0
    torch.cuda.synchronize()
    # the drop-in: empty COO arrays and an N x N CSC matrix without entries
    E, B, mapping = tet_mass_stiffness_matrices(f, er, ur, np.arange(T))
    assert E.size == 0 and B.size == 0
    S = mapping.to_csc(E)
    assert S.shape == (N, N) and S.nnz == 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_output_views_at_16_byte_offsets(wr90_field, precision):
    """Outputs need 16-byte alignment only: views 16 bytes into larger buffers get the
    values of freshly allocated outputs and leave the elements around them untouched; a
    view 8 bytes in (complex64) is refused before any launch."""
    import torch
    from emerge_amd._lib import EmError
    from emerge_amd.assembly import Ned2Plan
    f = wr90_field
    T = f.n_tets
    rng = np.random.default_rng(14)
    er, ur = ref_tensor(T, "scalar", rng), ref_tensor(T, "diag", rng)
    plan = Ned2Plan.from_field(f)
    n = plan.nnz_coo
    A = plan.assemble_a(er, ur, K0, precision)
    K, M = plan.assemble_km(er, ur, precision)
    shift = 1 if precision == "fp64" else 2  # elements in 16 bytes
    fill = 7.0 + 7.0j

    def padded():
        return torch.full((n + 2 * shift,), fill, dtype=A.dtype, device="cuda")

    bufs = [padded() for _ in range(3)]
    views = [b[shift:shift + n] for b in bufs]
    assert all(v.data_ptr() % 16 == 0 and v.data_ptr() % 32 != 0 for v in views)
    plan.assemble_a(er, ur, K0, precision, out=views[0])
    plan.assemble_km(er, ur, precision, out=(views[1], views[2]))
    assert torch.equal(views[0], A)
    assert torch.equal(views[1], K) and torch.equal(views[2], M)
    for b in bufs:
        assert bool((b[:shift] == fill).all()) and bool((b[shift + n:] == fill).all())
    if precision == "mixed":
        with pytest.raises(EmError, match="aligned"):
            plan.assemble_a(er, ur, K0, precision, out=padded()[1:1 + n])
