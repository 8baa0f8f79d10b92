"""GPU parity of em_material_prep (materials.hip, emerge_amd.materials) against the oracle
restatement of assembler.py:439-469, and the prepared tensors through the drop-in assembly
with the conductor tets left out."""
import numpy as np
import pytest

from oracle import materials as om
from tests.conftest import rel_fro

pytestmark = pytest.mark.gpu

F = 9.3e9


def _inputs(T, seed):
    rng = np.random.default_rng(seed)
    er = rng.uniform(1, 10, (3, 3, T)) + 1j * rng.normal(size=(3, 3, T))
    tand = rng.uniform(0, 0.02, (3, 3, T)) + 1j * rng.normal(size=(3, 3, T)) * 1e-3
    cond = rng.uniform(0, 50, (3, 3, T)) * (rng.random((3, 3, T)) < 0.3) + 0j
    # conductors and boundary values in cond[0,0]
    n = min(T, 40)
    pick = rng.choice(T, n, replace=False)
    edge = [1e4, 1e4 + 1j, 1e4 - 1j, 2e8, 1e8, 1e8 + 1j, 5e7, 1j, -1j, 3e6]
    cond[0, 0, pick] = np.resize(edge, n)
    return er, tand, cond


def _close(got, ref, er, tand, cond):
    """Within a few roundings of the terms' size: the GPU follows NumPy's fused
    complex products; a NumPy build without them differs by that much, at most."""
    scale = np.abs(er) * (1 + np.abs(tand)) + np.abs(cond) / (2 * np.pi * F * om.EPS0)
    return bool(np.all(np.abs(got - ref) <= 4 * 2.0 ** -52 * scale))


@pytest.mark.parametrize("T", [0, 1, 7, 20000, 700_001])
def test_material_prep_matches_oracle(T):
    import torch
    from emerge_amd.materials import prepare_materials
    er, tand, cond = _inputs(max(T, 1), 5)
    er, tand, cond = er[:, :, :T], tand[:, :, :T], cond[:, :, :T]
    ref, ids_ref, fd_ref = om.prepare_materials(er, tand, cond, F)
    got, ids, fd = prepare_materials(er, tand, cond, F)
    assert got.shape == (3, 3, T) and got.is_cuda
    assert _close(got.cpu().numpy(), ref, er, tand, cond)
    assert ids.dtype == np.int64 and np.array_equal(ids, ids_ref)
    assert fd == fd_ref
    # in place, on device tensors; PEC-only conductivities are not frequency dependent
    e = torch.as_tensor(np.ascontiguousarray(er), device="cuda")
    c2 = np.where(np.abs(cond) > 0, 2e9 + 0j, 0j)
    ref2, ids2, fd2 = om.prepare_materials(er, tand, c2, F)
    out, ids_b, fd_b = prepare_materials(e, tand, c2, F, out=e)
    assert out.data_ptr() == e.data_ptr()
    assert _close(e.cpu().numpy(), ref2, er, tand, c2)
    assert np.array_equal(ids_b, ids2) and fd_b == fd2 and not fd_b


def test_bad_arguments():
    import torch
    from emerge_amd.materials import prepare_materials
    er, tand, cond = _inputs(10, 6)
    with pytest.raises(ValueError):
        prepare_materials(er, tand[:, :, :9], cond, F)
    with pytest.raises(ValueError):
        prepare_materials(er, tand, cond, F,
                          out=torch.empty((3, 3, 10), dtype=torch.complex64, device="cuda"))


def test_prepared_materials_through_the_drop_in(wr90_field):
    """assemble_freq_matrix's flow (assembler.py:439-494): prepared er, conductor tets
    removed, then tet_mass_stiffness_matrices and to_csc(E - K0^2 B); equal to the oracle
    on the same prepared tensors and conductor list."""
    import oracle
    from emerge_amd.curlcurl import tet_mass_stiffness_matrices
    from emerge_amd.materials import prepare_materials
    f = wr90_field
    T = f.n_tets
    rng = np.random.default_rng(7)
    er = np.zeros((3, 3, T), complex)
    ur = np.zeros((3, 3, T), complex)
    tand = np.zeros((3, 3, T), complex)
    cond = np.zeros((3, 3, T), complex)
    for d in range(3):
        er[d, d] = rng.uniform(1, 4, T)
        ur[d, d] = 1.0
        tand[d, d] = rng.uniform(0, 0.01, T)
        cond[d, d] = np.where(np.arange(T) % 97 == 0, 5.8e7, 0.0)  # copper tets
    ref, ids_ref, fd_ref = om.prepare_materials(er, tand, cond, F)
    got, ids, fd = prepare_materials(er, tand, cond, F)
    assert np.array_equal(ids, ids_ref) and ids.size == len(range(0, T, 97)) and fd and fd_ref
    E, B, cmap = tet_mass_stiffness_matrices(f, got.cpu().numpy(), ur, ids)
    assert E.shape == (400 * (T - ids.size),)
    k0 = 2 * np.pi * F / 299792458.0
    A = cmap.to_csc(E - B * k0 ** 2)
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(f, ref, ur, ids)
    A_ref = oracle.CSCMap(rows, cols, f.n_field).to_csc(dE - dB * k0 ** 2)
    assert np.array_equal(A.indptr, A_ref.indptr) and np.array_equal(A.indices, A_ref.indices)
    assert rel_fro(A.data, A_ref.data) <= 1e-12


def test_volume_assembler_flow_and_caching(wr90_field):
    """emerge_amd.assembler.VolumeAssembler follows assemble_freq_matrix (assembler.py:
    423-494): A equals the oracle's (prepared materials, conductor tets removed, CSC of
    E - K0^2 B) at each frequency. The element matrices are reused while nothing is
    frequency dependent (a lossy dielectric, or PEC-only conductors) and rebuilt when a
    finite conductivity is present or caching is off."""
    import oracle
    from emerge_amd.assembler import VolumeAssembler
    f = wr90_field
    T = f.n_tets
    rng = np.random.default_rng(8)
    er = np.zeros((3, 3, T), complex)
    ur = np.zeros((3, 3, T), complex)
    tand = np.zeros((3, 3, T), complex)
    eps, td = rng.uniform(1, 4, T), rng.uniform(0, 0.01, T)
    for d in range(3):
        er[d, d], tand[d, d], ur[d, d] = eps, td, 1.0
    copper, pec = np.zeros_like(er), np.zeros_like(er)
    for d in range(3):
        copper[d, d] = np.where(np.arange(T) % 97 == 0, 5.8e7, 0.0)
        pec[d, d] = np.where(np.arange(T) % 89 == 0, 1e9, 0.0)

    def oracle_A(cond, freq):
        ref, ids, fd = om.prepare_materials(er, tand, cond, freq)
        dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(f, ref, ur, ids)
        k0 = 2 * np.pi * freq / 299792458
        return oracle.CSCMap(rows, cols, f.n_field).to_csc(dE - dB * k0 ** 2), ids, fd

    for cond, dep in ((np.zeros_like(er), False), (pec, False), (copper, True)):
        va = VolumeAssembler(f)
        for i, freq in enumerate((8.2e9, 9.7e9, 11.1e9)):
            r = va.assemble_freq_matrix(er, ur, tand, cond, freq)
            A_ref, ids, fd = oracle_A(cond, freq)
            assert r.is_frequency_dependent == fd == dep
            assert np.array_equal(r.conductor_tets, ids)
            A = r.to_scipy()
            assert np.array_equal(A.indptr, A_ref.indptr)
            assert np.array_equal(A.indices, A_ref.indices)
            assert rel_fro(A.data, A_ref.data) <= 1e-12, (dep, freq)
            # frequency dependent (copper's finite conductivity): the first call assembles
            # A, the second finds the same arrays and builds the affine decomposition,
            # the third only combines
            assert va.assemblies == (min(i + 1, 2) if dep else 1)
        va.assemble_freq_matrix(er, ur, tand, cond, 10e9, cache_matrices=False)
        assert va.assemblies == (3 if dep else 2)
    # mixed precision through the same flow
    a64 = VolumeAssembler(f).assemble_freq_matrix(er, ur, tand, copper, 10e9).data
    a32 = VolumeAssembler(f, "mixed").assemble_freq_matrix(er, ur, tand, copper, 10e9).data
    assert a32.dtype.itemsize == 8
    assert rel_fro(a32.cpu().numpy().astype(complex), a64.cpu().numpy()) < 1e-6


def test_volume_assembler_scattering_shares_the_cache(wr90_field):
    """assemble_scattering_matrix's volume step (assembler.py:808-868) is the frequency
    path's on the same caches: equal results bit for bit, element matrices built once and
    reused by either entry point while nothing is frequency dependent, and the oracle's A."""
    import torch
    import oracle
    from emerge_amd.assembler import VolumeAssembler
    f = wr90_field
    T = f.n_tets
    rng = np.random.default_rng(9)
    er = np.zeros((3, 3, T), complex)
    ur = np.zeros((3, 3, T), complex)
    tand = np.zeros((3, 3, T), complex)
    cond = np.zeros((3, 3, T), complex)
    for d in range(3):
        er[d, d], tand[d, d], ur[d, d] = rng.uniform(1, 4, T), rng.uniform(0, 0.01, T), 1.0
        cond[d, d] = np.where(np.arange(T) % 89 == 0, 1e9, 0.0)  # PEC-only conductors
    va = VolumeAssembler(f)
    s1 = va.assemble_scattering_matrix(er, ur, tand, cond, 9.1e9)
    f1 = va.assemble_freq_matrix(er, ur, tand, cond, 9.1e9)
    s2 = va.assemble_scattering_matrix(er, ur, tand, cond, 10.3e9)
    assert va.assemblies == 1
    assert torch.equal(s1.data, f1.data)
    assert np.array_equal(s1.conductor_tets, f1.conductor_tets)
    fresh = VolumeAssembler(f).assemble_freq_matrix(er, ur, tand, cond, 10.3e9)
    assert torch.equal(s2.data, fresh.data)
    ref, ids, _ = om.prepare_materials(er, tand, cond, 10.3e9)
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(f, ref, ur, ids)
    k0 = 2 * np.pi * 10.3e9 / 299792458
    A_ref = oracle.CSCMap(rows, cols, f.n_field).to_csc(dE - dB * k0 ** 2)
    A = s2.to_scipy()
    assert np.array_equal(A.indptr, A_ref.indptr) and np.array_equal(A.indices, A_ref.indices)
    assert rel_fro(A.data, A_ref.data) <= 1e-12
    va.assemble_scattering_matrix(er, ur, tand, cond, 10.3e9, cache_matrices=False)
    assert va.assemblies == 2


def test_volume_assembler_eig_matrices(wr90_field):
    """assemble_eig_matrix's volume step (assembler.py:1115-1153): K and M converted to CSC
    separately, conductor tets removed; equal to the oracle's."""
    import oracle
    from emerge_amd.assembler import VolumeAssembler
    f = wr90_field
    T = f.n_tets
    rng = np.random.default_rng(9)
    er = np.zeros((3, 3, T), complex)
    ur = np.zeros((3, 3, T), complex)
    tand = np.zeros((3, 3, T), complex)
    cond = np.zeros((3, 3, T), complex)
    for d in range(3):
        er[d, d] = rng.uniform(1, 4, T) * (1 + 0.2 * d)
        ur[d, d] = rng.uniform(1, 2, T)
        tand[d, d] = rng.uniform(0, 0.01, T)
        cond[d, d] = np.where(np.arange(T) % 53 == 0, 1e9, 0.0)
    K, M, cm, ids = VolumeAssembler(f).assemble_eig_matrices(er, ur, tand, cond, 9e9)
    ref, ids_ref, _ = om.prepare_materials(er, tand, cond, 9e9)
    assert np.array_equal(ids, ids_ref) and ids.size > 0
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(f, ref, ur, ids_ref)
    cmap = oracle.CSCMap(rows, cols, f.n_field)
    ex = cm.export("indptr", "indices")
    assert np.array_equal(ex["indptr"].cpu().numpy(), cmap.indptr)
    assert np.array_equal(ex["indices"].cpu().numpy(), cmap.indices)
    assert rel_fro(K.cpu().numpy(), cmap.to_csc_data(dE)) <= 1e-12
    assert rel_fro(M.cpu().numpy(), cmap.to_csc_data(dB)) <= 1e-12


def _device_inputs(T, seed):
    """er, ur, tand, cond as complex (3, 3, T) arrays: lossy anisotropic er, a full
    non-symmetric ur, a few copper tets."""
    rng = np.random.default_rng(seed)
    er = rng.normal(size=(3, 3, T)) * 0.2 + 0j
    ur = rng.normal(size=(3, 3, T)) * 0.3 + 1j * rng.normal(size=(3, 3, T)) * 0.05
    tand = np.zeros((3, 3, T), complex)
    cond = np.zeros((3, 3, T), complex)
    for d in range(3):
        er[d, d] += rng.uniform(2, 4, T)
        ur[d, d] += 2.0
        tand[d, d] = rng.uniform(0, 0.01, T)
        cond[d, d] = np.where(np.arange(T) % 41 == 5, 5.8e7, 0.0)
    return er, ur, tand, cond


@pytest.mark.parametrize("which", ["wr90", "three_tets"])
def test_volume_assembler_device_inputs(wr90_field, which):
    """er, ur, tand and cond all handed over as CUDA (3, 3, T) tensors, as _prepare's
    docstring allows: assemble_freq_matrix and assemble_eig_matrices equal the oracle on
    the same arrays. The three-tet mesh is the case where a (3, 3, T) device tensor has the
    shape of a tet-major (T, 3, 3) one."""
    import torch

    import oracle
    from emerge_amd.assembler import VolumeAssembler
    from emerge_amd.mesh import Nedelec2Field, TetMesh, hex_box_mesh
    if which == "wr90":
        f = wr90_field
    else:
        hm = hex_box_mesh(1e-2, 1e-2, 1e-2, 1, 1, 1)
        f = Nedelec2Field(TetMesh(hm.nodes, hm.tets[:, :3]))
    T = f.n_tets
    er, ur, tand, cond = _device_inputs(T, 12)
    dev = [torch.as_tensor(np.ascontiguousarray(a), device="cuda") for a in (er, ur, tand, cond)]
    freq = 9.1e9
    k0 = 2 * np.pi * freq / 299792458
    ref, ids, fd = om.prepare_materials(er, tand, cond, freq)
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(f, ref, ur, ids)
    cmap = oracle.CSCMap(rows, cols, f.n_field)
    r = VolumeAssembler(f).assemble_freq_matrix(*dev, freq)
    assert np.array_equal(r.conductor_tets, ids) and r.is_frequency_dependent == fd
    A = r.to_scipy()
    assert np.array_equal(A.indptr, cmap.indptr) and np.array_equal(A.indices, cmap.indices)
    assert rel_fro(A.data, cmap.to_csc_data(dE - dB * k0 ** 2)) <= 1e-12
    K, M, cm, ids_e = VolumeAssembler(f).assemble_eig_matrices(*dev, freq)
    assert np.array_equal(ids_e, ids)
    assert rel_fro(K.cpu().numpy(), cmap.to_csc_data(dE)) <= 1e-12
    assert rel_fro(M.cpu().numpy(), cmap.to_csc_data(dB)) <= 1e-12


def test_volume_assembler_cached_blocks_keep_their_symmetry(wr90_field):
    """Cached K/M built from full tensors, reused by a call with diagonal tensors of
    another frequency (nothing frequency dependent, same plan): the result is the oracle's
    A of the materials the blocks were built from, through the mirrored CSC gathers that
    every element block allows (the blocks are symmetric whatever the tensors)."""
    import oracle
    from emerge_amd.assembler import VolumeAssembler
    f = wr90_field
    T = f.n_tets
    er, ur, tand, cond = _device_inputs(T, 13)
    cond[:] = 0
    freq1, freq2 = 8.7e9, 10.3e9
    va = VolumeAssembler(f)
    va.assemble_freq_matrix(er, ur, tand, cond, freq1)
    # same plan (no conductors), nothing frequency dependent: the blocks are reused
    er_d = np.zeros_like(er)
    ur_d = np.zeros_like(ur)
    for d in range(3):
        er_d[d, d], ur_d[d, d] = er[d, d], ur[d, d]
    r = va.assemble_freq_matrix(er_d, ur_d, tand, cond, freq2)
    assert va.assemblies == 1
    ref, ids, _ = om.prepare_materials(er, tand, cond, freq1)
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(f, ref, ur, ids)
    k0 = 2 * np.pi * freq2 / 299792458
    want = oracle.CSCMap(rows, cols, f.n_field).to_csc_data(dE - dB * k0 ** 2)
    assert rel_fro(r.data.cpu().numpy(), want) <= 1e-12


def test_volume_assembler_detects_structure_of_device_inputs(wr90_field):
    """Diagonal (3, 3, T) inputs handed over as CUDA tensors are recognised on the device
    (diagonal er', diagonal ur: the diagonal kernel and the mirrored gather) and give the
    oracle's A; with a finite conductivity the frequency-dependent path (A formed in the
    element kernel, one gather) matches the oracle and the K/M combine."""
    import torch

    import oracle
    from emerge_amd import _lib
    from emerge_amd.assembler import VolumeAssembler
    f = wr90_field
    T = f.n_tets
    rng = np.random.default_rng(21)
    er = np.zeros((3, 3, T), complex)
    ur = np.zeros((3, 3, T), complex)
    tand = np.zeros((3, 3, T), complex)
    cond = np.zeros((3, 3, T), complex)
    for d in range(3):
        er[d, d] = rng.uniform(1, 4, T) * (1 + 0.1 * d)
        ur[d, d] = rng.uniform(1, 2, T)
        tand[d, d] = rng.uniform(0, 0.01, T)
        cond[d, d] = np.where(np.arange(T) % 31 == 0, 2e3, 0.0)  # lossy, not a conductor
    dev = [torch.as_tensor(a, device="cuda") for a in (er, ur, tand, cond)]
    freq = 9.4e9
    k0 = 2 * np.pi * freq / 299792458
    va = VolumeAssembler(f)
    plan, er_m, ur_m, ids, fd = va._prepare(*dev, freq)
    assert er_m.kind == _lib.EM_MAT_DIAG and ur_m.kind == _lib.EM_MAT_DIAG and fd
    r = va.assemble_freq_matrix(*dev, freq)
    assert r.is_frequency_dependent and va.cached_matrices is None and va.assemblies == 1
    ref, ids_ref, _ = om.prepare_materials(er, tand, cond, freq)
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(f, ref, ur, ids_ref)
    want = oracle.CSCMap(rows, cols, f.n_field).to_csc_data(dE - dB * k0 ** 2)
    assert rel_fro(r.data.cpu().numpy(), want) <= 1e-12
    K, M = plan.assemble_km(er_m, ur_m, "fp64")
    km = r.cscmap.to_csc_combine(K, M, k0, block_symmetric=True)
    assert rel_fro(r.data.cpu().numpy(), km.cpu().numpy()) <= 1e-14
    # a second frequency rebuilds (nothing is cached for frequency-dependent materials)
    va.assemble_freq_matrix(*dev, 10.1e9)
    assert va.assemblies == 2


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_volume_assembler_lossy_medium_affine_path(wr90_field, precision):
    """A lossy medium swept in frequency (finite conductivity, arrays unchanged): from the
    second frequency on A = K - K0^2 M(er (1 - j tand)) + j K0^2/(W0 EPS0) M(cond) from
    three cached CSC arrays; each frequency equals the oracle's full reassembly, and the
    direct path within rounding. A changed array (tand) leaves the affine path and
    assembles afresh."""
    import torch

    import oracle
    from emerge_amd.assembler import VolumeAssembler
    f = wr90_field
    T = f.n_tets
    rng = np.random.default_rng(33)
    er = np.zeros((3, 3, T), complex)
    ur = np.zeros((3, 3, T), complex)
    tand = np.zeros((3, 3, T), complex)
    cond = np.zeros((3, 3, T), complex)
    for d in range(3):
        er[d, d] = rng.uniform(1, 4, T) * (1 + 0.1 * d)
        ur[d, d] = rng.uniform(1, 2, T)
        tand[d, d] = rng.uniform(0, 0.01, T)
        cond[d, d] = np.where(np.arange(T) % 7 == 0, rng.uniform(1, 5e3, T), 0.0)
        cond[d, d] = np.where(np.arange(T) % 101 == 3, 1e9, cond[d, d])  # PEC tets too
    dev = [torch.as_tensor(a, device="cuda") for a in (er, ur, tand, cond)]
    tol = 1e-12 if precision == "fp64" else 1e-6
    va = VolumeAssembler(f, precision)
    for i, freq in enumerate((8.3e9, 9.1e9, 10.7e9, 11.9e9)):
        r = va.assemble_freq_matrix(*dev, freq)
        assert r.is_frequency_dependent
        ref, ids, fd = om.prepare_materials(er, tand, cond, freq)
        assert fd and np.array_equal(r.conductor_tets, ids) and ids.size > 0
        dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(f, ref, ur, ids)
        k0 = 2 * np.pi * freq / 299792458
        want = oracle.CSCMap(rows, cols, f.n_field).to_csc_data(dE - dB * k0 ** 2)
        assert rel_fro(r.data.cpu().numpy(), want) <= tol, (i, freq)
        assert va.assemblies == min(i + 1, 2)
        if i == 2:
            # the affine path against this frequency's own reassembly
            fresh = VolumeAssembler(f, precision).assemble_freq_matrix(*dev, freq).data
            assert rel_fro(r.data.cpu().numpy(), fresh.cpu().numpy()) <= \
                (1e-13 if precision == "fp64" else 1e-6)
    # another loss tangent: the arrays differ, so the call reassembles
    dev[2] = dev[2] * 1.5
    r = va.assemble_freq_matrix(*dev, 10e9)
    assert va.assemblies == 3
    ref, ids, _ = om.prepare_materials(er, tand * 1.5, cond, 10e9)
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(f, ref, ur, ids)
    k0 = 2 * np.pi * 10e9 / 299792458
    want = oracle.CSCMap(rows, cols, f.n_field).to_csc_data(dE - dB * k0 ** 2)
    assert rel_fro(r.data.cpu().numpy(), want) <= tol
