"""CPU tests of the host side: the C-ABI library loads and exports every declared symbol,
argument validation fails with the documented status codes before touching a GPU, the
material descriptors, and the mesh / DOF tables against a loop restatement of
mesh3d.py:707-774 and nedelec2.py:65-80."""
import ctypes
import os
import re

import numpy as np
import pytest

from emerge_amd import _lib
from emerge_amd.mesh import (DOFCODES_NED2, EDGE_SLOTS, FACE_SLOTS, MeshError, Nedelec2Field,
                             TetMesh, delaunay_box_mesh, hex_box_mesh)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "emerge_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\s*\*)\s+(em_\w+)\s*\(", src, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _declared_functions()
    assert len(declared) >= 14
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) == set(_lib.EXPORTED)
    assert b"gfx950" in L.em_version()


def test_plan_create_rejects_bad_arguments_without_gpu():
    L = _lib.lib()
    h = ctypes.c_void_p()
    f = Nedelec2Field(hex_box_mesh(1, 1, 1, 1, 1, 1))
    m = f.mesh
    keep = []

    def p(a, dt=np.int64):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data_as(ctypes.c_void_p)

    args = [ctypes.byref(h), 0, p(m.nodes, np.float64), m.n_nodes, p(m.tets), m.n_tets,
            p(m.tris), m.n_tris, p(m.edges), m.n_edges, p(f.tet_to_field), p(m.tet_to_edge),
            p(m.tet_to_tri), f.n_field, None, 0]
    # unsupported DOF set (the reference raises ValueError on unknown codes)
    bad_codes = DOFCODES_NED2.copy()
    bad_codes[0] = 65
    st = L.em_ned2_plan_create(*args, p(bad_codes), 20)
    assert st == _lib.EM_EUNSUPPORTED
    assert b"SECOND_MIXED_VOLAKIS" in L.em_last_error()
    st = L.em_ned2_plan_create(*args, p(DOFCODES_NED2[:6]), 6)
    assert st == _lib.EM_EUNSUPPORTED
    # null tables
    st = L.em_ned2_plan_create(ctypes.byref(h), 0, None, 1, None, 0, None, 0, None, 0, None, None,
                               None, 1, None, 0, p(DOFCODES_NED2), 20)
    assert st == _lib.EM_EINVAL
    # node id out of range
    tets = m.tets.copy()
    tets[0, 0] = m.n_nodes + 5
    a2 = list(args)
    a2[4] = p(tets)
    assert L.em_ned2_plan_create(*a2, p(DOFCODES_NED2), 20) == _lib.EM_EINVAL
    # repeated node in a tet
    tets = m.tets.copy()
    tets[1, 0] = tets[0, 0]
    a2[4] = p(tets)
    assert L.em_ned2_plan_create(*a2, p(DOFCODES_NED2), 20) == _lib.EM_EINVAL
    # DOF id out of range
    t2f = f.tet_to_field.copy()
    t2f[3, 2] = f.n_field
    a3 = list(args)
    a3[10] = p(t2f)
    assert L.em_ned2_plan_create(*a3, p(DOFCODES_NED2), 20) == _lib.EM_EINVAL
    # null handles everywhere else
    assert L.em_ned2_assemble_km(None, None, None, None, None, 0, None) == _lib.EM_EINVAL
    assert L.em_csc_map_sizes(None, None, None, None) == _lib.EM_EINVAL
    assert L.em_ned2_plan_destroy(None) == _lib.EM_OK


def test_material_prep_rejects_bad_arguments_without_gpu():
    """em_material_prep checks its arguments before any device call: a negative tet
    count, missing pointers and a non-positive omega (the reference divides by zero at
    f = 0) give EM_EINVAL."""
    L = _lib.lib()
    n = ctypes.c_int64(-5)
    fd = ctypes.c_int32(-5)
    fake = ctypes.c_void_p(4096)  # never dereferenced: refused first
    st = L.em_material_prep(fake, fake, fake, 10, 0.0, 1e8, 1e4, fake, fake, ctypes.byref(n),
                            ctypes.byref(fd), None)
    assert st == _lib.EM_EINVAL and b"omega" in L.em_last_error()
    st = L.em_material_prep(fake, fake, fake, -1, 1.0, 1e8, 1e4, fake, fake, ctypes.byref(n),
                            ctypes.byref(fd), None)
    assert st == _lib.EM_EINVAL
    st = L.em_material_prep(fake, None, fake, 10, 1.0, 1e8, 1e4, fake, fake, ctypes.byref(n),
                            ctypes.byref(fd), None)
    assert st == _lib.EM_EINVAL
    st = L.em_material_prep(fake, fake, fake, 10, 1.0, 1e8, 1e4, fake, fake, None,
                            ctypes.byref(fd), None)
    assert st == _lib.EM_EINVAL and n.value == -5 and fd.value == -5


def test_material_prep_noalloc_refuses_beyond_its_scratch():
    """em_material_prep_noalloc's fixed scratch covers 16,777,216 tets; one more is refused
    with EM_EINVAL before any device call, as are the arguments em_material_prep refuses."""
    L = _lib.lib()
    n = ctypes.c_int64(-5)
    fd = ctypes.c_int32(-5)
    fake = ctypes.c_void_p(4096)  # never dereferenced: refused first
    st = L.em_material_prep_noalloc(fake, fake, fake, (1 << 24) + 1, 1.0, 1e8, 1e4, fake, fake,
                                    ctypes.byref(n), ctypes.byref(fd), None)
    assert st == _lib.EM_EINVAL and b"16,777,216" in L.em_last_error()
    st = L.em_material_prep_noalloc(fake, fake, fake, 10, 0.0, 1e8, 1e4, fake, fake,
                                    ctypes.byref(n), ctypes.byref(fd), None)
    assert st == _lib.EM_EINVAL and b"omega" in L.em_last_error()
    assert n.value == -5 and fd.value == -5


def test_status_to_exception():
    with pytest.raises(_lib.UnsupportedError):
        _lib.check(_lib.EM_EUNSUPPORTED)
    with pytest.raises(_lib.DegenerateTetError):
        _lib.check(_lib.EM_EDEGENERATE)
    _lib.check(_lib.EM_OK)


def test_material_descriptors():
    import torch
    from emerge_amd.assembly import material
    T = 5
    cpu = torch.device("cpu")
    rng = np.random.default_rng(0)
    full = rng.normal(size=(3, 3, T)) + 1j * rng.normal(size=(3, 3, T))
    m = material(full, T, cpu)
    assert (m.kind, m.tet_stride, m.comp_stride) == (_lib.EM_MAT_FULL, 1, T)
    diag = np.zeros((3, 3, T), complex)
    for d in range(3):
        diag[d, d] = rng.normal(size=T)
    m = material(diag, T, cpu)
    assert (m.kind, m.tet_stride, m.comp_stride) == (_lib.EM_MAT_DIAG, 1, 4 * T)
    # the in-place diagonal read: value(t, d) = data[t + d * 4T]
    flat = m.tensor.numpy().ravel()
    for d in range(3):
        assert np.array_equal(flat[np.arange(T) + d * 4 * T], diag[d, d])
    sc = np.zeros((3, 3, T), complex)
    for d in range(3):
        sc[d, d] = 2 - 1j
    m = material(sc, T, cpu)
    assert (m.kind, m.tet_stride) == (_lib.EM_MAT_SCALAR, 1)
    m = material(3.0, T, cpu)
    assert (m.kind, m.tet_stride) == (_lib.EM_MAT_SCALAR, 0)
    m = material(torch.ones(T, 3, dtype=torch.complex128), T, cpu)
    assert (m.kind, m.tet_stride, m.comp_stride) == (_lib.EM_MAT_DIAG, 3, 1)
    m = material(torch.ones(T, 3, 3, dtype=torch.complex128), T, cpu)
    assert (m.kind, m.tet_stride, m.comp_stride) == (_lib.EM_MAT_FULL, 9, 1)
    with pytest.raises(ValueError):
        material(np.ones((3, 3, T + 1), complex), T, cpu)
    with pytest.raises(ValueError):
        material(torch.ones(T + 1), T, cpu)


def _loop_tables(tets):
    """Loop restatement of Mesh3D._pre_update's edge/tri extraction and slot tables
    (mesh3d.py:707-774) with Python sets and dicts, for small meshes."""
    edgeset, triset = set(), set()
    for t in range(tets.shape[1]):
        i1, i2, i3, i4 = sorted(int(x) for x in tets[:, t])
        edgeset |= {(i1, i2), (i1, i3), (i1, i4), (i2, i3), (i2, i4), (i3, i4)}
        triset |= {(i1, i2, i3), (i1, i2, i4), (i1, i3, i4), (i2, i3, i4)}
    edges = np.array(sorted(edgeset)).T
    tris = np.array(sorted(triset)).T
    inv_e = {tuple(edges[:, i]): i for i in range(edges.shape[1])}
    inv_t = {tuple(tris[:, i]): i for i in range(tris.shape[1])}
    idset1 = ((1, 2), (1, 3), (1, 4), (2, 3), (4, 2), (3, 4))
    idset2 = ((1, 2, 3), (1, 3, 4), (1, 4, 2), (2, 3, 4))
    t2e = np.zeros((6, tets.shape[1]), int)
    t2t = np.zeros((4, tets.shape[1]), int)
    for it in range(tets.shape[1]):
        t = tets[:, it]
        t2e[:, it] = [inv_e[tuple(sorted((int(t[i - 1]), int(t[j - 1]))))] for i, j in idset1]
        t2t[:, it] = [inv_t[tuple(sorted((int(t[i - 1]), int(t[j - 1]), int(t[k - 1]))))]
                      for i, j, k in idset2]
    return edges, tris, t2e, t2t


@pytest.mark.parametrize("mk", ["hex", "delaunay"])
def test_mesh_tables_equal_loop_restatement(mk):
    mesh = hex_box_mesh(1, 2, 3, 2, 2, 3) if mk == "hex" else delaunay_box_mesh(300, seed=9)
    edges, tris, t2e, t2t = _loop_tables(mesh.tets)
    assert np.array_equal(mesh.edges, edges)
    assert np.array_equal(mesh.tris, tris)
    assert np.array_equal(mesh.tet_to_edge, t2e)
    assert np.array_equal(mesh.tet_to_tri, t2t)
    f = Nedelec2Field(mesh)
    E, F = mesh.n_edges, mesh.n_tris
    # nedelec2.py:65-80 layout
    assert np.array_equal(f.tet_to_field[0:6], t2e)
    assert np.array_equal(f.tet_to_field[6:12], t2e + E)
    assert np.array_equal(f.tet_to_field[12:16], t2t + 2 * E)
    assert np.array_equal(f.tet_to_field[16:20], t2t + 2 * E + F)
    assert f.n_field == 2 * E + 2 * F
    f.diagnose()
    assert EDGE_SLOTS[4] == (3, 1) and FACE_SLOTS[2] == (0, 3, 1)


def test_waveguide_mesh_counts(wr90_field):
    m = wr90_field.mesh
    assert (m.n_tets, m.n_nodes, m.n_edges, m.n_tris, wr90_field.n_field) == \
        (7200, 1650, 9669, 15220, 49778)
    assert (m.signed_dets() > 0).all()


def test_delaunay_generator_oriented_and_conforming():
    mesh = delaunay_box_mesh(3000, seed=5)
    assert (mesh.signed_dets() > 0).all()
    cnt = np.bincount(mesh.tet_to_tri.ravel(), minlength=mesh.n_tris)
    assert cnt.min() >= 1 and cnt.max() <= 2
    # same seed, same mesh
    again = delaunay_box_mesh(3000, seed=5)
    assert np.array_equal(again.tets, mesh.tets) and np.array_equal(again.nodes, mesh.nodes)


def test_mesh_errors():
    with pytest.raises(MeshError):
        TetMesh(np.zeros((3, 4)), np.array([[0], [1], [2], [2]]))
    with pytest.raises(MeshError):
        TetMesh(np.zeros((3, 4)), np.array([[0], [1], [2], [4]]))


def test_prepared_material_structure_matches_the_arithmetic():
    """The structure the volume assembler assigns to er(1 - j tand) - j cond/(w eps0)
    from its host inputs is the structure of the oracle's result: scalar only when every
    component on the diagonal is equal, diagonal when no off-diagonal entry is nonzero
    (an off-diagonal loss tangent alone keeps it diagonal), full otherwise."""
    from emerge_amd.assembler import _prepared_kind
    from oracle import materials as om
    rng = np.random.default_rng(4)
    T = 50
    off = ([0, 0, 1, 1, 2, 2], [1, 2, 0, 2, 0, 1])

    def diag(vals):
        a = np.zeros((3, 3, T), complex)
        for d in range(3):
            a[d, d] = vals[d]
        return a

    s = rng.uniform(1, 5, T) + 0j
    cases = {
        "scalar": (diag([s, s, s]), diag([s / 100] * 3), diag([s * 0] * 3)),
        "diag_er": (diag([s, s * 2, s]), diag([s / 100] * 3), diag([s * 0] * 3)),
        "diag_cond": (diag([s, s, s]), diag([s / 100] * 3), diag([s, s * 0, s])),
        "offdiag_tand": (diag([s, s, s]), diag([s / 100] * 3) + 0.01 * (1 - np.eye(3))[:, :, None],
                         diag([s * 0] * 3)),
        "full_er": (diag([s, s, s]) + 0.1 * (1 - np.eye(3))[:, :, None], diag([s / 100] * 3),
                    diag([s * 0] * 3)),
        "full_cond": (diag([s, s, s]), diag([s * 0] * 3),
                      np.where(np.eye(3)[:, :, None] > 0, 0j, 3.0 + 0j) * np.ones((3, 3, T))),
    }
    want = {"scalar": _lib.EM_MAT_SCALAR, "diag_er": _lib.EM_MAT_DIAG,
            "diag_cond": _lib.EM_MAT_DIAG, "offdiag_tand": _lib.EM_MAT_SCALAR,
            "full_er": _lib.EM_MAT_FULL, "full_cond": _lib.EM_MAT_FULL}
    for name, (er, tand, cond) in cases.items():
        kind = _prepared_kind(er, tand, cond)
        assert kind == want[name], name
        out, _, _ = om.prepare_materials(er, tand, cond, 10e9)
        if kind != _lib.EM_MAT_FULL:
            assert not np.any(out[off]), name
        if kind == _lib.EM_MAT_SCALAR:
            assert np.array_equal(out[0, 0], out[1, 1]) and np.array_equal(out[0, 0], out[2, 2])
        if kind == _lib.EM_MAT_DIAG:
            assert not (np.array_equal(out[0, 0], out[1, 1]) and
                        np.array_equal(out[0, 0], out[2, 2])), name
    # torch tensors are inspected where they live (here on the host): same answers
    import torch
    for name, (er, tand, cond) in cases.items():
        assert _prepared_kind(torch.from_numpy(er), torch.from_numpy(tand), cond) == \
            want[name], name
