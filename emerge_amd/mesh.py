"""Tetrahedral meshes, their edge/face tables and the Nedelec-2 DOF map.

The tables are the ones the reference's assembler reads from its mesh object:

* ``edges`` (2, E) / ``tris`` (3, F): every edge and face of the mesh as a tuple of
  node ids sorted ascending, the tuples themselves in lexicographic order
  (reference ``src/emerge/_emerge/mesh3d.py:707-732``).
* ``tet_to_edge`` (6, T) / ``tet_to_tri`` (4, T): slot tables. Slot ``e`` of a tet is
  the edge between stored vertex positions ``EDGE_SLOTS[e]``; slot ``f`` is the face
  ``FACE_SLOTS[f]`` (``mesh3d.py:757-774``; ``_idset1``/``_idset2`` there are 1-based).
* ``tet_to_field`` (20, T) and ``n_field``: the DOF map of the second-order mixed
  (Volakis) Nedelec space, DOF codes ``[64]*6 + [66]*6 + [128]*4 + [129]*4``
  (``elements/nedelec2.py:37-80``, ``elements/dofsets.py:30-48, 114-115``).

The reference builds these with Python loops over tets; here they are built with
sorted-key numpy operations, which give the same arrays.

The generators make seeded synthetic meshes: the hex-split box of the waveguide
parity case and Delaunay meshes of a box. Every generated tet is re-oriented so that
``det[v1-v0, v2-v0, v3-v0] > 0`` (the orientation gmsh emits and the reference relies on).
"""
from __future__ import annotations

import numpy as np

# Stored-vertex positions of the 6 edge slots and 4 face slots of a tet
# (reference mesh3d.py:760-761, converted to 0-based positions).
EDGE_SLOTS = ((0, 1), (0, 2), (0, 3), (1, 2), (3, 1), (2, 3))
FACE_SLOTS = ((0, 1, 2), (0, 2, 3), (0, 3, 1), (1, 2, 3))

# DOF codes of ElementSpace.SECOND_MIXED_VOLAKIS: edge functions 0 and 2, face
# functions 0 and 1 (reference dofsets.py:39, 114-115).
DOFCODES_NED2 = np.array([64] * 6 + [66] * 6 + [128] * 4 + [129] * 4, dtype=np.int64)


MISSING_ID = -1234  # reference mesh3d.py:29

class MeshError(ValueError):
    """Raised for malformed or degenerate meshes (reference: ``MeshException``)."""


def _unique_rows_sorted(keys_cols: list[np.ndarray], n_nodes: int):
    """Lexicographically sorted unique tuples and the inverse map.

    ``keys_cols`` are equal-length int64 columns (tuple components, each tuple already
    sorted ascending). Packs the tuple into one int64 key when that cannot overflow.
    """
    k = len(keys_cols)
    if n_nodes ** k < 2 ** 62:
        key = keys_cols[0].astype(np.int64)
        for c in keys_cols[1:]:
            key = key * n_nodes + c
        uniq, inv = np.unique(key, return_inverse=True)
        cols = []
        rem = uniq.copy()
        for _ in range(k):
            cols.append(rem % n_nodes)
            rem //= n_nodes
        return np.stack(cols[::-1]), inv.astype(np.int64)
    stacked = np.ascontiguousarray(np.stack(keys_cols, axis=1))
    uniq, inv = np.unique(stacked, axis=0, return_inverse=True)
    return uniq.T.copy(), inv.reshape(-1).astype(np.int64)


class TetMesh:
    """Nodes, tets and the derived edge/face tables of a tetrahedral mesh.

    Attribute names follow the reference ``Mesh3D`` so that code written against it
    (``field.mesh.tets``, ``field.mesh.tet_to_edge`` ...) reads the same arrays.
    """

    def __init__(self, nodes: np.ndarray, tets: np.ndarray):
        nodes = np.ascontiguousarray(np.asarray(nodes, dtype=np.float64))
        tets = np.ascontiguousarray(np.asarray(tets, dtype=np.int64))
        if nodes.ndim != 2 or nodes.shape[0] != 3:
            raise MeshError(f"nodes must have shape (3, Nn), got {nodes.shape}")
        if tets.ndim != 2 or tets.shape[0] != 4:
            raise MeshError(f"tets must have shape (4, T), got {tets.shape}")
        nn = nodes.shape[1]
        if tets.size and (tets.min() < 0 or tets.max() >= nn):
            raise MeshError("tet node id out of range")
        self.nodes = nodes
        self.tets = tets
        self._build_tables()

    def _build_tables(self) -> None:
        tets = self.tets
        nn = self.nodes.shape[1]
        T = tets.shape[1]
        st = np.sort(tets, axis=0)
        if T and np.any(st[1:] == st[:-1]):
            raise MeshError("tet with repeated node id")
        # --- edges: sorted pairs, lexicographic (mesh3d.py:714-731)
        ea = [st[a] for a, b in ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))]
        eb = [st[b] for a, b in ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))]
        self.edges, _ = _unique_rows_sorted([np.concatenate(ea), np.concatenate(eb)], nn)
        # --- tris: sorted triples, lexicographic
        tri_local = ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3))
        cols = [np.concatenate([st[t[c]] for t in tri_local]) for c in range(3)]
        self.tris, _ = _unique_rows_sorted(cols, nn)
        # --- slot tables (mesh3d.py:757-774): slot tuples sorted, then looked up
        self.tet_to_edge = self._lookup_slots(EDGE_SLOTS, self.edges, nn)
        self.tet_to_tri = self._lookup_slots(FACE_SLOTS, self.tris, nn)

    def _lookup_slots(self, slots, table: np.ndarray, nn: int) -> np.ndarray:
        k = table.shape[0]
        T = self.tets.shape[1]
        out = np.empty((len(slots), T), dtype=np.int64)
        packable = nn ** k < 2 ** 62
        if packable:
            tkey = table[0].astype(np.int64)
            for c in range(1, k):
                tkey = tkey * nn + table[c]
        for s, pos in enumerate(slots):
            v = np.sort(np.stack([self.tets[p] for p in pos]), axis=0)
            if packable:
                key = v[0].astype(np.int64)
                for c in range(1, k):
                    key = key * nn + v[c]
                out[s] = np.searchsorted(tkey, key)
            else:
                rec = np.rec.fromarrays(list(v))
                trec = np.rec.fromarrays(list(table))
                out[s] = np.searchsorted(trec, rec)
        return out

    # reference Mesh3D naming
    @property
    def n_nodes(self) -> int:
        return self.nodes.shape[1]

    @property
    def n_tets(self) -> int:
        return self.tets.shape[1]

    @property
    def n_edges(self) -> int:
        return self.edges.shape[1]

    @property
    def n_tris(self) -> int:
        return self.tris.shape[1]

    def signed_dets(self) -> np.ndarray:
        """det[v1-v0, v2-v0, v3-v0] per tet in stored vertex order."""
        return tet_dets(self.nodes, self.tets)

    @property
    def tri_to_edge(self) -> np.ndarray:
        """(3, F): edges (t0,t1), (t1,t2), (t0,t2) of each face's sorted node tuple
        (reference mesh3d.py:832-839)."""
        if getattr(self, "_tri_to_edge", None) is None:
            nn = self.n_nodes
            ekey = self.edges[0] * nn + self.edges[1]
            t = self.tris
            self._tri_to_edge = np.stack([np.searchsorted(ekey, t[a] * nn + t[b])
                                          for a, b in ((0, 1), (1, 2), (0, 2))])
        return self._tri_to_edge

    @property
    def tri_to_tet(self) -> np.ndarray:
        """(2, F): the tets sharing each face, lower tet id first; ``MISSING_ID`` in row 1
        for boundary faces (reference mesh3d.py:780-795)."""
        if getattr(self, "_tri_to_tet", None) is None:
            flat = self.tet_to_tri.ravel(order="F")
            tet = np.repeat(np.arange(self.n_tets, dtype=np.int64), 4)
            srt = np.argsort(flat, kind="stable")
            ft, tt = flat[srt], tet[srt]
            _, first, counts = np.unique(ft, return_index=True, return_counts=True)
            out = np.full((2, self.n_tris), MISSING_ID, dtype=np.int64)
            out[0, ft[first]] = tt[first]
            two = counts == 2
            out[1, ft[first[two]]] = tt[first[two] + 1]
            self._tri_to_tet = out
        return self._tri_to_tet

    def boundary_tris(self) -> np.ndarray:
        """Ids of the faces that belong to one tet only (the outer surface)."""
        return np.flatnonzero(self.tri_to_tet[1] == MISSING_ID)


class Nedelec2Field:
    """The DOF map of the second-order mixed Nedelec space on a ``TetMesh``.

    Mirrors the attributes of the reference ``Nedelec2`` that the assembly reads
    (``elements/nedelec2.py:37-101``): ``mesh``, ``tet_to_field``, ``n_field``,
    ``dofcodes3d``, ``get_tet_to_field()``.

    Global numbering: code 64 on edge e -> e; code 66 -> e + E; code 128 on face f ->
    f + 2E; code 129 -> f + F + 2E (``nedelec2.py:71-78``).
    """

    def __init__(self, mesh: TetMesh):
        self.mesh = mesh
        E, F = mesh.n_edges, mesh.n_tris
        self.nedges = E
        self.ntris = F
        self.ntets = mesh.n_tets
        self.n_tets = mesh.n_tets
        self.n_edges = E
        self.n_tris = F
        self.nedof = 2
        self.nfdof = 2
        self.n_field = 2 * E + 2 * F
        self.dofcodes3d = DOFCODES_NED2.copy()
        t2f = np.empty((20, mesh.n_tets), dtype=np.int64)
        t2f[0:6] = mesh.tet_to_edge
        t2f[6:12] = mesh.tet_to_edge + E
        t2f[12:16] = mesh.tet_to_tri + 2 * E
        t2f[16:20] = mesh.tet_to_tri + F + 2 * E
        self.tet_to_field = t2f
        self.n_tet_dofs = 20
        self.n_tri_dofs = 8

    @property
    def edge_to_field(self) -> np.ndarray:
        """(2, E): the two DOFs of each edge, e and e + E (nedelec2.py:70-73)."""
        E = self.n_edges
        return np.stack([np.arange(E), np.arange(E) + E]).astype(np.int64)

    @property
    def tri_to_field(self) -> np.ndarray:
        """(8, F): the DOFs of each face's closure, rows 0-2 its edges' first DOFs, 3-5
        their second DOFs, 6-7 the two face DOFs (nedelec2.py:70-79)."""
        E, F = self.n_edges, self.n_tris
        t2e = self.mesh.tri_to_edge
        f = np.arange(F, dtype=np.int64)
        return np.concatenate([t2e, t2e + E, (f + 2 * E)[None], (f + F + 2 * E)[None]])

    def get_tet_to_field(self) -> np.ndarray:
        return self.tet_to_field

    def diagnose(self) -> None:
        """Every DOF is touched by some tet (reference ``nedelec2.py:103-107``)."""
        seen = np.zeros(self.n_field, dtype=bool)
        seen[self.tet_to_field.ravel()] = True
        if not seen.all():
            raise MeshError("DOF not touched by any tet")


# ---------------------------------------------------------------------------
# geometry helpers
# ---------------------------------------------------------------------------

def tet_dets(nodes: np.ndarray, tets: np.ndarray) -> np.ndarray:
    p = [nodes[:, tets[k]] for k in range(4)]
    a, b, c = p[1] - p[0], p[2] - p[0], p[3] - p[0]
    return np.einsum("ij,ij->j", a, np.cross(b, c, axis=0))


def orient_positive(nodes: np.ndarray, tets: np.ndarray) -> np.ndarray:
    """Swap the last two stored vertices of every tet with negative determinant."""
    tets = tets.copy()
    neg = tet_dets(nodes, tets) < 0
    tmp = tets[2, neg].copy()
    tets[2, neg] = tets[3, neg]
    tets[3, neg] = tmp
    return tets


def _morton_order(pts01: np.ndarray, bits: int = 10) -> np.ndarray:
    """Order of points along a Morton (Z-order) curve; ``pts01`` in [0,1]^3, shape (n,3)."""
    q = np.clip((pts01 * (1 << bits)).astype(np.int64), 0, (1 << bits) - 1)
    code = np.zeros(q.shape[0], dtype=np.int64)
    for b in range(bits):
        for d in range(3):
            code |= ((q[:, d] >> b) & 1) << (3 * b + d)
    return np.argsort(code, kind="stable")


def _sort_tets_by_nodes(tets: np.ndarray) -> np.ndarray:
    """Order tets lexicographically by their sorted node ids (memory locality)."""
    st = np.sort(tets, axis=0)
    order = np.lexsort((st[3], st[2], st[1], st[0]))
    return tets[:, order]


# ---------------------------------------------------------------------------
# generators
# ---------------------------------------------------------------------------

def hex_box_mesh(lx: float, ly: float, lz: float, nx: int, ny: int, nz: int) -> TetMesh:
    """Box [0,lx]x[0,ly]x[0,lz] split into nx*ny*nz hexes, each cut into 6 tets.

    All hexes use the same main diagonal (corner (0,0,0) to (1,1,1)), so the split is
    conforming. Tet k of a hex follows the axis permutation path from the low corner to
    the high corner. Tets are re-oriented to positive determinant.
    """
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    zs = np.linspace(0.0, lz, nz + 1)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    nodes = np.stack([X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")])

    def nid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    I, J, K = I.ravel(order="F"), J.ravel(order="F"), K.ravel(order="F")
    unit = np.eye(3, dtype=np.int64)
    perms = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
    tets = []
    for p in perms:
        c1 = unit[p[0]]
        c2 = unit[p[0]] + unit[p[1]]
        v0 = nid(I, J, K)
        v1 = nid(I + c1[0], J + c1[1], K + c1[2])
        v2 = nid(I + c2[0], J + c2[1], K + c2[2])
        v3 = nid(I + 1, J + 1, K + 1)
        tets.append(np.stack([v0, v1, v2, v3]))
    # hex-major order: the 6 tets of one hex are adjacent
    tets = np.stack(tets, axis=2).reshape(4, -1)
    tets = orient_positive(nodes, tets)
    return TetMesh(nodes, tets)


def delaunay_box_mesh(n_points: int, seed: int, box=(1.0, 1.0, 1.0),
                      boundary_per_side: int | None = None, jitter: float = 0.3,
                      rel_vol_tol: float = 1e-9) -> TetMesh:
    """Seeded Delaunay mesh of a box: uniform interior points plus a jittered grid on
    each face (points stay on their face) and the 8 corners.

    Points are ordered along a Morton curve and tets lexicographically by node ids, so
    neighbouring tets share nearby DOF numbers. Tets whose volume is below
    ``rel_vol_tol`` times the mean are dropped; the rest are oriented positive.
    """
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    box = np.asarray(box, dtype=np.float64)
    if boundary_per_side is None:
        boundary_per_side = max(2, int(round(n_points ** (1.0 / 3.0))))
    m = boundary_per_side
    pts = [rng.random((n_points, 3))]
    g = (np.arange(m) + 0.5) / m
    U, V = np.meshgrid(g, g, indexing="ij")
    U, V = U.ravel(), V.ravel()
    for axis in range(3):
        a1, a2 = [d for d in range(3) if d != axis]
        for side in (0.0, 1.0):
            ju = U + (rng.random(U.size) - 0.5) * (jitter / m)
            jv = V + (rng.random(V.size) - 0.5) * (jitter / m)
            p = np.empty((U.size, 3))
            p[:, axis] = side
            p[:, a1] = ju
            p[:, a2] = jv
            pts.append(p)
    corners = np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)], dtype=np.float64)
    pts.append(corners)
    pts01 = np.concatenate(pts)
    pts01 = pts01[_morton_order(pts01)]
    tri = Delaunay(pts01)
    tets = tri.simplices.T.astype(np.int64)
    nodes = (pts01 * box).T.copy()
    det = tet_dets(nodes, tets)
    keep = np.abs(det) > rel_vol_tol * np.mean(np.abs(det))
    tets = tets[:, keep]
    # drop nodes no tet uses (points qhull left out), renumber preserving order
    used = np.zeros(nodes.shape[1], dtype=bool)
    used[tets.ravel()] = True
    if not used.all():
        remap = np.cumsum(used) - 1
        nodes = nodes[:, used]
        tets = remap[tets]
    tets = orient_positive(nodes, tets)
    tets = _sort_tets_by_nodes(tets)
    return TetMesh(nodes, tets)


def delaunay_mesh_with_tets(target_tets: int, seed: int, box=(1.0, 1.0, 1.0)) -> TetMesh:
    """Delaunay box mesh sized for about ``target_tets`` tets (about 6.5 tets per point)."""
    n = max(8, int(target_tets / 6.55))
    return delaunay_box_mesh(n, seed, box=box)


class Legrange2Field:
    """The DOF map of second-order Lagrange elements on a ``TetMesh`` (reference
    ``Legrange2``, elements/leg2.py:29-60): node DOFs first (the node ids), then one DOF
    per edge (edge id + n_nodes); ``tet_to_field`` (10, T) lists a tet's 4 vertices in
    stored order and its 6 edges in slot order."""

    def __init__(self, mesh: TetMesh):
        self.mesh = mesh
        self.nnodes = mesh.n_nodes
        self.nedges = mesh.n_edges
        self.ntets = mesh.n_tets
        self.n_tets = mesh.n_tets
        self.n_field = mesh.n_nodes + mesh.n_edges
        t2f = np.empty((10, mesh.n_tets), dtype=np.int64)
        t2f[:4] = mesh.tets
        t2f[4:] = mesh.tet_to_edge + mesh.n_nodes
        self.tet_to_field = t2f
        self.n_tet_dofs = 10
