"""GPU CSC map (precompute_csc_pattern / to_csc, csc_cast.py:65-177) against the oracle's
restatement and against SciPy's COO->CSC conversion; and the drop-in entry point
(curlcurl.py:222-270 + assembler.py:494) end to end."""
import time

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from tests.conftest import ref_tensor, rel_fro

pytestmark = pytest.mark.gpu


def _field_plan(field, tetids=None):
    from emerge_amd.assembly import Ned2Plan
    return Ned2Plan.from_field(field, tetids=tetids)


@pytest.mark.parametrize("which", ["wr90", "delaunay"])
def test_csc_pattern_bit_exact(which, wr90_field, delaunay_field):
    from emerge_amd.assembly import CscMap
    field = wr90_field if which == "wr90" else delaunay_field
    T = field.n_tets
    I = ref_tensor(T, "identity", None)
    _, _, rows, cols = oracle.tet_mass_stiffness_matrices(field, I, I, np.array([]))
    ref = oracle.CSCMap(rows, cols, field.n_field)
    plan = _field_plan(field)
    cm = CscMap(plan)
    assert cm.nnz == ref.nnz and cm.N == field.n_field
    ex = {k: v.cpu().numpy() for k, v in
          cm.export("indptr", "indices", "perm", "target_csc", "col_offsets").items()}
    assert np.array_equal(ex["indptr"], ref.indptr)
    assert np.array_equal(ex["indices"], ref.indices)
    assert np.array_equal(ex["col_offsets"], ref.col_offsets)
    assert np.array_equal(ex["perm"], ref.perm)
    assert np.array_equal(ex["target_csc"], ref.target_csc)
    # SciPy: same pattern
    S = sp.coo_matrix((np.ones(rows.shape[0]), (rows, cols)),
                      shape=(field.n_field, field.n_field)).tocsc()
    S.sum_duplicates()
    S.sort_indices()
    assert np.array_equal(S.indptr, ex["indptr"]) and np.array_equal(S.indices, ex["indices"])


def test_to_csc_values_and_combine(delaunay_field):
    import torch
    from emerge_amd.assembly import CscMap
    rng = np.random.default_rng(21)
    T = delaunay_field.n_tets
    er = ref_tensor(T, "scalar", rng)
    ur = ref_tensor(T, "diag", rng)
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(delaunay_field, er, ur, np.array([]),
                                                            local_frame=True)
    k0 = 2 * np.pi * 10e9 / 299792458.0
    N = delaunay_field.n_field
    S = sp.coo_matrix((dE - dB * k0 ** 2, (rows, cols)), shape=(N, N)).tocsc()
    S.sum_duplicates()
    S.sort_indices()
    plan = _field_plan(delaunay_field)
    cm = CscMap(plan)
    K, M = plan.assemble_km(er, ur, "fp64")
    v1 = cm.to_csc_combine(K, M, k0).cpu().numpy()
    A = plan.assemble_a(er, ur, k0, "fp64")
    v2 = cm.to_csc(A).cpu().numpy()
    assert rel_fro(v1, S.data) <= 1e-12
    assert rel_fro(v2, S.data) <= 1e-12
    # oracle's own scatter restatement on the oracle's COO values
    ref = oracle.CSCMap(rows, cols, N)
    assert rel_fro(ref.to_csc_data(dE - dB * k0 ** 2), S.data) <= 1e-14
    # mixed precision: complex64 COO summed in FP64
    A32 = plan.assemble_a(er, ur, k0, "mixed")
    v3 = cm.to_csc(A32)
    assert v3.dtype == torch.complex64
    assert rel_fro(v3.cpu().numpy().astype(np.complex128), S.data) <= 1e-6


def test_dropin_matches_reference_pipeline(wr90_field):
    """tet_mass_stiffness_matrices + cscmap.to_csc(Evec - Bvec*K0**2), as the reference's
    Assembler.assemble_freq_matrix calls them (assembler.py:485-494)."""
    from emerge_amd.curlcurl import tet_mass_stiffness_matrices
    rng = np.random.default_rng(3)
    T = wr90_field.n_tets
    er = ref_tensor(T, "scalar", rng)
    ur = ref_tensor(T, "identity", rng)
    conductor = np.array([5, 17, 400, 7000])
    Evec, Bvec, cscmap = tet_mass_stiffness_matrices(wr90_field, er, ur, conductor)
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(wr90_field, er, ur, conductor)
    assert Evec.dtype == np.complex128 and Evec.shape == dE.shape
    assert rel_fro(Evec, dE) <= 1e-12 and rel_fro(Bvec, dB) <= 1e-12
    K0 = 2 * np.pi * 10e9 / 299792458.0
    A = cscmap.to_csc(Evec - Bvec * K0 ** 2)
    ref = oracle.CSCMap(rows, cols, wr90_field.n_field).to_csc(dE - dB * K0 ** 2)
    assert isinstance(A, sp.csc_matrix)
    assert np.array_equal(A.indptr, ref.indptr) and np.array_equal(A.indices, ref.indices)
    assert rel_fro(A.data, ref.data) <= 1e-12
    # second call reuses the map (the reference caches it, assembler.py:490-491)
    E2, B2, c2 = tet_mass_stiffness_matrices(wr90_field, er, ur, conductor, cscmap)
    assert c2 is cscmap and np.array_equal(E2, Evec)


@pytest.mark.parametrize("kinds", [("scalar", "diag"), ("diag", "diag"), ("scalar", "scalar")])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_blocks_symmetric_and_mirrored_gather(delaunay_field, kinds, precision):
    """Scalar and diagonal tensors give bit-symmetric 20x20 blocks, so the mirrored-read
    gather (EM_CSC_BLOCK_SYMMETRIC) returns exactly the plain gather's values, for
    to_csc and the fused combine."""
    import torch
    from emerge_amd.assembly import CscMap
    rng = np.random.default_rng(22)
    T = delaunay_field.n_tets
    er, ur = ref_tensor(T, kinds[0], rng), ref_tensor(T, kinds[1], rng)
    plan = _field_plan(delaunay_field)
    cm = CscMap(plan)
    k0 = 2 * np.pi * 9.5e9 / 299792458.0
    K, M = plan.assemble_km(er, ur, precision)
    A = plan.assemble_a(er, ur, k0, precision)
    for X in (K, M, A):
        blk = X.view(T, 20, 20)
        assert torch.equal(blk, blk.transpose(1, 2))
    assert torch.equal(cm.to_csc(A), cm.to_csc(A, block_symmetric=True))
    assert torch.equal(cm.to_csc_combine(K, M, k0),
                       cm.to_csc_combine(K, M, k0, block_symmetric=True))


@pytest.mark.parametrize("precision", ["fp64", "mixed", "fp64_ref"])
def test_full_tensor_blocks_bit_symmetric(delaunay_field, precision):
    """The reference computes each element matrix for i <= j only and mirrors it
    (curlcurl.py:357-371), so its 20x20 blocks are symmetric whatever the tensors are;
    the kernels do the same: with full non-symmetric complex eps_r and mu_r every block
    of K, M, A and of a multi-wavenumber A equals its transpose bit for bit, and the
    mirrored-read gather returns the plain gather's values exactly."""
    import torch
    from emerge_amd.assembly import CscMap
    rng = np.random.default_rng(24)
    T = delaunay_field.n_tets
    er, ur = ref_tensor(T, "full", rng), ref_tensor(T, "full", rng)
    assert not np.allclose(er, er.transpose(1, 0, 2))
    plan = _field_plan(delaunay_field)
    cm = CscMap(plan)
    k0 = 2 * np.pi * 9.5e9 / 299792458.0
    K, M = plan.assemble_km(er, ur, precision)
    A = plan.assemble_a(er, ur, k0, precision)
    mats = [K, M, A] + list(plan.assemble_a(er, ur, [0.5 * k0, k0, 1.5 * k0], precision))
    for X in mats:
        blk = X.view(T, 20, 20)
        assert torch.equal(blk, blk.transpose(1, 2))
        assert float(torch.linalg.norm(blk.real)) > 0
    for X in (K, M, A):
        assert torch.equal(cm.to_csc(X), cm.to_csc(X, block_symmetric=True))


def test_mirrored_gather_symmetric_full_tensors(delaunay_field):
    """Symmetric full tensors (the full-tensor kernel path, whose antisymmetric metric
    part is then rounding noise): symmetric blocks, and the mirrored gather agrees with
    the plain one."""
    import torch
    from emerge_amd.assembly import CscMap
    rng = np.random.default_rng(23)
    T = delaunay_field.n_tets
    er, ur = ref_tensor(T, "sym", rng), ref_tensor(T, "sym", rng)
    plan = _field_plan(delaunay_field)
    cm = CscMap(plan)
    A = plan.assemble_a(er, ur, 2 * np.pi * 9.5e9 / 299792458.0, "fp64")
    blk = A.view(T, 20, 20)
    assert float(torch.linalg.norm(blk - blk.transpose(1, 2)) / torch.linalg.norm(blk)) < 1e-15
    a, b = cm.to_csc(A), cm.to_csc(A, block_symmetric=True)
    assert float(torch.linalg.norm(a - b) / torch.linalg.norm(a)) < 1e-15


def _fan_field(n):
    """n tets around one edge (bottom-top axis, ring of n nodes): the axis edge's DOF
    columns hold more than 256 distinct rows for n = 30."""
    from emerge_amd.mesh import Nedelec2Field, TetMesh, orient_positive
    ang = 2 * np.pi * np.arange(n) / n
    nodes = np.concatenate([[[0.0, 0.0, 0.0], [0.0, 0.0, 1.0]],
                            np.stack([np.cos(ang), np.sin(ang), np.full(n, 0.5)], 1)]).T
    tets = np.array([[0, 1, 2 + i, 2 + (i + 1) % n] for i in range(n)]).T
    return Nedelec2Field(TetMesh(nodes, orient_positive(nodes, tets)))


@pytest.mark.parametrize("mesh", ["delaunay", "fan30"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_incidence_gather_equals_segment_gather(delaunay_field, mesh, precision):
    """The per-column incidence gather (default for full maps) and the per-entry perm/seg
    gather give the same values bit for bit (same contributors, same order), plain and
    mirrored, to_csc and combine; long columns (> 256 entries) go through LDS windows."""
    import torch
    from emerge_amd.assembly import CscMap
    field = delaunay_field if mesh == "delaunay" else _fan_field(30)
    rng = np.random.default_rng(24)
    T = field.n_tets
    er, ur = ref_tensor(T, "scalar", rng), ref_tensor(T, "diag", rng)
    plan = _field_plan(field)
    cm = CscMap(plan)
    lens = np.diff(cm.export("indptr")["indptr"].cpu().numpy())
    if mesh == "fan30":
        assert lens.max() > 256
    k0 = 2 * np.pi * 8e9 / 299792458.0
    K, M = plan.assemble_km(er, ur, precision)
    for sym in (False, True):
        a = cm.to_csc(K, block_symmetric=sym)
        b = cm.to_csc(K, block_symmetric=sym, segment_gather=True)
        assert torch.equal(a, b), sym
        a = cm.to_csc_combine(K, M, k0, block_symmetric=sym)
        b = cm.to_csc_combine(K, M, k0, block_symmetric=sym, segment_gather=True)
        assert torch.equal(a, b), sym
    # and against the oracle's CSC of the same mesh
    if precision == "fp64":
        dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(field, er, ur, np.array([]),
                                                                local_frame=True)
        ref = oracle.CSCMap(rows, cols, field.n_field).to_csc_data(dE - k0 ** 2 * dB)
        assert rel_fro(cm.to_csc_combine(K, M, k0).cpu().numpy(), ref) <= 1e-12


@pytest.mark.parametrize("mesh", ["delaunay", "fan30"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_staged_gather_non_symmetric_blocks(delaunay_field, mesh, precision):
    """COO values whose blocks are not symmetric (random data: the assembly itself mirrors
    its upper triangle, curlcurl.py:358-376, so its blocks are): the staged gather
    (contributions written block-transposed in FP64, then the mirrored column sum) equals
    the plain incidence gather and the segment gather bit for bit, for to_csc and the
    combine, on the full map and on a restricted (free-DOF) map."""
    import torch
    from emerge_amd.assembly import CscMap
    from emerge_amd.pec import pec_solve_ids
    field = delaunay_field if mesh == "delaunay" else _fan_field(30)
    plan = _field_plan(field)
    cm = CscMap(plan)
    k0 = 2 * np.pi * 9e9 / 299792458.0
    dt = torch.complex128 if precision == "fp64" else torch.complex64
    g = torch.Generator().manual_seed(31)
    K = torch.randn(cm.nnz_coo, generator=g, dtype=dt).cuda()
    M = torch.randn(cm.nnz_coo, generator=g, dtype=dt).cuda()
    work = torch.empty(cm.nnz_coo, dtype=torch.complex128, device="cuda")
    ref = cm.to_csc_combine(K, M, k0, segment_gather=True)
    assert torch.equal(cm.to_csc_staged(K, M, k0, work=work), ref)
    assert torch.equal(cm.to_csc_combine(K, M, k0), ref)
    assert torch.equal(cm.to_csc_staged(K, None, 0.0, work=work), cm.to_csc(K, segment_gather=True))
    # restricted map: the free-DOF block of an all-PEC box
    sel = pec_solve_ids(plan, field, [], field.mesh.boundary_tris())
    sub = cm.restrict(sel)
    assert torch.equal(sub.to_csc_staged(K, M, k0, work=work), sub.to_csc_combine(K, M, k0))


@pytest.mark.parametrize("n", [102, 103, 614, 615, 1200, 3276])
def test_long_columns_and_the_limit(n):
    """A DOF in n tets gives a column of 20 n COO entries. Up to 102 tets (2,040 entries)
    the column kernel sorts them in LDS; from 103 on the long-column kernel ranks them,
    with the rows in LDS up to 614 tets (12,280 of 12,288 keys) and in global scratch
    from 615 on. Either way the pattern, perm and target_csc equal the oracle's and every
    gather matches its CSC; a DOF in 3,277 tets (more than the 65,535 entries the uint16
    row positions index) is refused with EM_EUNSUPPORTED (the reference's CPU argsort has
    no such limit, csc_cast.py:65-160)."""
    import torch
    from emerge_amd._lib import UnsupportedError
    from emerge_amd.assembly import CscMap
    field = _fan_field(n)
    rng = np.random.default_rng(25)
    T = field.n_tets
    er, ur = ref_tensor(T, "scalar", rng), ref_tensor(T, "diag", rng)
    plan = _field_plan(field)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cm = CscMap(plan)
    torch.cuda.synchronize()
    print(f"map of a {n}-tet fan ({20 * n} entries in the axis columns): "
          f"{(time.perf_counter() - t0) * 1e3:.1f} ms", flush=True)
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(field, er, ur, np.array([]),
                                                            local_frame=True)
    ref = oracle.CSCMap(rows, cols, field.n_field)
    ex = {k: v.cpu().numpy() for k, v in
          cm.export("indptr", "indices", "perm", "target_csc").items()}
    assert np.diff(ex["indptr"]).max() > 256
    for k in ("indptr", "indices", "perm", "target_csc"):
        assert np.array_equal(ex[k], getattr(ref, k)), k
    k0 = 2 * np.pi * 8e9 / 299792458.0
    K, M = plan.assemble_km(er, ur, "fp64")
    want = ref.to_csc_data(dE - k0 ** 2 * dB)
    got = cm.to_csc_combine(K, M, k0)
    assert rel_fro(got.cpu().numpy(), want) <= 1e-12
    assert torch.equal(got, cm.to_csc_combine(K, M, k0, block_symmetric=True))
    assert torch.equal(got, cm.to_csc_combine(K, M, k0, segment_gather=True))
    if n == 1200:
        with pytest.raises(UnsupportedError, match="3277 tets \\(at most 3276"):
            CscMap(_field_plan(_fan_field(3277)))


@pytest.mark.parametrize("case", ["scalar-diag", "full-full", "not-symmetric"])
def test_dropin_to_csc_picks_a_gather_with_the_plain_result(wr90_field, case):
    """CSCMapping.to_csc checks the blocks of the data it is handed. The assembly's blocks
    are always symmetric -- the reference fills i <= j and mirrors (curlcurl.py:358-376),
    full non-symmetric tensors included -- and so is E - K0^2 B: those go to the mirrored
    gather. Data with blocks that are not symmetric (built by the caller) go to the staged
    gather. Either way the values equal the plain per-column gather's bit for bit and the
    oracle's to_csc of the same data within 1e-12."""
    import torch
    from emerge_amd.curlcurl import tet_mass_stiffness_matrices
    rng = np.random.default_rng(23)
    T = wr90_field.n_tets
    kinds = ("full", "full") if case == "full-full" else ("scalar", "diag")
    er, ur = ref_tensor(T, kinds[0], rng), ref_tensor(T, kinds[1], rng)
    Evec, Bvec, cscmap = tet_mass_stiffness_matrices(wr90_field, er, ur, np.array([]))
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(wr90_field, er, ur, np.array([]))
    K0 = 2 * np.pi * 9.5e9 / 299792458.0
    data, ref_data = Evec - Bvec * K0 ** 2, dE - dB * K0 ** 2
    if case == "not-symmetric":
        upper = np.triu(np.ones((20, 20)), 1).astype(bool).ravel()
        scale = np.where(np.tile(upper, T), 1.5, 1.0)
        data, ref_data = data * scale, ref_data * scale
    blocks = data.reshape(-1, 20, 20)
    assert np.array_equal(blocks, blocks.transpose(0, 2, 1)) == (case != "not-symmetric")
    A = cscmap.to_csc(data)
    plain = cscmap._dev.to_csc(torch.from_numpy(data).cuda()).cpu().numpy()
    assert np.array_equal(A.data, plain)
    ref = oracle.CSCMap(rows, cols, wr90_field.n_field).to_csc(ref_data)
    assert rel_fro(A.data, ref.data) <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kinds", [("scalar", "diag"), ("full", "full")])
def test_direct_csc_assembly_equals_coo_path(delaunay_field, precision, kinds):
    """em_ned2_assemble_csc (CSC values without the COO arrays) equals to_csc of the COO
    values bit for bit: A against to_csc(assemble_a), K and M against to_csc of
    assemble_km; for symmetric and non-symmetric blocks; on a plan with conductor tets
    removed and on the PEC-restricted map of that plan; and A against the oracle."""
    import torch
    from emerge_amd.assembly import CscMap, Ned2Plan
    from emerge_amd.curlcurl import plan_for
    from emerge_amd.pec import pec_solve_ids
    f = delaunay_field
    T = f.n_tets
    rng = np.random.default_rng(31)
    er = ref_tensor(T, kinds[0], rng)
    ur = ref_tensor(T, kinds[1], rng, complex_=False)
    k0 = 2 * np.pi * 9.7e9 / 299792458.0
    cond = np.sort(rng.choice(T, 60, replace=False))
    for plan in (Ned2Plan.from_field(f), plan_for(f, cond)):
        cm = CscMap(plan)
        e, m = plan.material(er), plan.material(ur)
        sym = kinds[0] != "full"
        A = plan.assemble_a(e, m, k0, precision)
        K, M = plan.assemble_km(e, m, precision)
        assert torch.equal(plan.assemble_csc(cm, e, m, k0, "A", precision),
                           cm.to_csc(A, block_symmetric=sym))
        assert torch.equal(plan.assemble_csc(cm, e, m, 0.0, "K", precision),
                           cm.to_csc(K, block_symmetric=sym))
        assert torch.equal(plan.assemble_csc(cm, e, m, 0.0, "M", precision),
                           cm.to_csc(M, block_symmetric=sym))
        sel = pec_solve_ids(plan, f, cond, f.mesh.boundary_tris())
        sub = cm.restrict(sel)
        assert torch.equal(plan.assemble_csc(sub, e, m, k0, "A", precision), sub.to_csc(A))
    # the last plan (conductors removed) against the oracle
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(f, er, ur, cond, local_frame=True)
    want = oracle.CSCMap(rows, cols, f.n_field).to_csc_data(dE - dB * k0 ** 2)
    got = plan.assemble_csc(cm, e, m, k0, "A", precision).cpu().numpy()
    assert rel_fro(got, want) <= (1e-12 if precision == "fp64" else 1e-6)
    # a map of another plan is refused
    with pytest.raises(RuntimeError):
        Ned2Plan.from_field(f).assemble_csc(cm, er, ur, k0, "A", precision)
