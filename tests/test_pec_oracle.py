"""CPU pins of the PEC-elimination oracle and of the face tables it reads."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import pec as opec


def _geometric_wall_dofs(field):
    """DOFs whose edge or face lies in one wall plane of the mesh's bounding box: on a
    box with every wall PEC these are exactly the prescribed DOFs."""
    m = field.mesh
    x = m.nodes
    lo, hi = x.min(axis=1), x.max(axis=1)
    tol = 1e-12 * (hi - lo).max()

    def on_wall(ids):  # ids (k, n): True where all k nodes share one wall coordinate
        res = np.zeros(ids.shape[1], bool)
        for d in range(3):
            for v in (lo[d], hi[d]):
                res |= np.all(np.abs(x[d][ids] - v) < tol, axis=0)
        return res

    E, F = m.n_edges, m.n_tris
    be = np.flatnonzero(on_wall(m.edges))
    bf = np.flatnonzero(on_wall(m.tris))
    return set(np.concatenate([be, be + E, bf + 2 * E, bf + 2 * E + F]).tolist()), bf


def test_tri_tables_match_bruteforce(wr90_field):
    from emerge_amd.mesh import Nedelec2Field, delaunay_box_mesh
    for f in (wr90_field, Nedelec2Field(delaunay_box_mesh(150, seed=5))):
        m = f.mesh
        assert np.array_equal(m.tri_to_edge, opec.tri_to_edge(m.tris, m.edges))
        assert np.array_equal(m.tri_to_tet, opec.tri_to_tet(m.tet_to_tri, m.n_tris))


def test_box_boundary_faces(wr90_field):
    m = wr90_field.mesh
    nx, ny, nz = 10, 5, 24
    # every boundary quad of the Kuhn split is two triangles
    assert len(m.boundary_tris()) == 4 * (nx * ny + ny * nz + nx * nz)
    _, bf = _geometric_wall_dofs(wr90_field)
    assert np.array_equal(np.sort(m.boundary_tris()), bf)


def test_field_face_tables_consistent(wr90_field):
    """tri_to_field columns agree with the tets' DOFs: face DOFs of slot k are rows 12+k
    and 16+k of tet_to_field, the face's edges are edges of the tet."""
    f = wr90_field
    m = f.mesh
    t2f, tri2f = f.tet_to_field, f.tri_to_field
    for k in range(4):
        faces = m.tet_to_tri[k]
        assert np.array_equal(tri2f[6, faces], t2f[12 + k])
        assert np.array_equal(tri2f[7, faces], t2f[16 + k])
        for r in range(3):
            e = tri2f[r, faces]
            assert np.all(np.any(m.tet_to_edge == e[None], axis=0))
            assert np.array_equal(tri2f[3 + r, faces], e + f.n_edges)
    assert np.array_equal(f.edge_to_field[1], f.edge_to_field[0] + f.n_edges)


def test_pec_dof_set_equals_walls(wr90_field):
    f = wr90_field
    m = f.mesh
    geo, bf = _geometric_wall_dofs(f)
    got = opec.pec_dof_set(f.tet_to_field, [], bf, m.tri_to_edge, f.edge_to_field,
                           f.tri_to_field)
    assert got == geo
    ids = opec.solve_ids(f.n_field, got)
    assert len(ids) == f.n_field - len(geo) and np.all(np.diff(ids) > 0)


def test_pec_dof_set_conductor_tets(wr90_field):
    f = wr90_field
    cond = [0, 5, 77, 7199]
    got = opec.pec_dof_set(f.tet_to_field, cond, [], f.mesh.tri_to_edge, f.edge_to_field,
                           f.tri_to_field)
    assert got == set(f.tet_to_field[:, cond].ravel().tolist())


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_select_block_matches_scipy(seed):
    rng = np.random.default_rng(seed)
    n = 300
    A = sp.random(n, n, density=0.03, random_state=seed, format="csc", dtype=np.complex128,
                  data_rvs=lambda k: rng.normal(size=k) + 1j * rng.normal(size=k))
    A = A.tolil()
    A[:, 7] = 0  # empty column
    A = A.tocsc()
    A.data[::11] = 0  # explicit zeros stay in the pattern
    A.sort_indices()
    keep = rng.random(n) < 0.7
    ids = np.flatnonzero(keep)
    ref = A[:, ids][ids, :].tocsc()
    ref.sort_indices()
    ip, ix, dx = opec.select_block(A.indptr, A.indices, A.data, ids, n)
    assert np.array_equal(ip, ref.indptr) and np.array_equal(ix, ref.indices)
    assert np.array_equal(dx, ref.data)
    # the eigen path's np.ix_ form is the same block
    ref2 = A[np.ix_(ids, ids)].tocsc()
    assert (ref2 != ref).nnz == 0
