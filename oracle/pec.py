"""ORACLE -- test infrastructure, not product code.

CPU restatement of the PEC elimination that follows the assembly in the reference:

* ``pec_dof_set`` -- ``Assembler.assemble_freq_matrix``'s PEC DOF collection
  (src/emerge/_emerge/physics/microwave/assembly/assembler.py:533-564), loop for loop:
  the ``tet_to_field`` column of every conductor tet, then for each PEC face the
  ``edge_to_field`` columns of its ``tri_to_edge`` edges and its ``tri_to_field`` column;
* ``solve_ids`` -- the ascending complement (assembler.py:742);
* ``select_block`` -- ``A[:, solve_ids][solve_ids, :]`` (src/emerge/_emerge/solver.py:2071)
  restated on the CSC arrays: keep the kept columns, inside each the kept rows (order
  preserved), renumber both;
* ``tri_to_edge`` / ``tri_to_tet`` -- brute-force restatements of mesh3d.py:780-795 and
  :832-839 with Python dictionaries, to pin the vectorised tables of ``emerge_amd.mesh``.

Pins (tests/test_pec_oracle.py): ``select_block`` equals SciPy's own slicing (the call
the reference makes) on random matrices with empty columns and explicit zeros;
``pec_dof_set`` equals the geometric set of DOFs whose edge or face lies in a wall of a
box whose walls are all PEC.
"""
from __future__ import annotations

import numpy as np

MISSING_ID = -1234  # mesh3d.py:29


def tri_to_edge(nodes_tris: np.ndarray, edges: np.ndarray) -> np.ndarray:
    """(3, F): edges (t0,t1), (t1,t2), (t0,t2) of each sorted face (mesh3d.py:832-839)."""
    look = {}
    for e in range(edges.shape[1]):
        a, b = int(edges[0, e]), int(edges[1, e])
        look[(a, b)] = e
        look[(b, a)] = e
    F = nodes_tris.shape[1]
    out = np.zeros((3, F), dtype=np.int64)
    for f in range(F):
        t = sorted(int(x) for x in nodes_tris[:, f])
        out[0, f] = look[(t[0], t[1])]
        out[1, f] = look[(t[1], t[2])]
        out[2, f] = look[(t[0], t[2])]
    return out


def tri_to_tet(tet_to_tri: np.ndarray, n_tris: int) -> np.ndarray:
    """(2, F): tets of each face in ascending tet order, MISSING_ID when there is one
    (mesh3d.py:780-795)."""
    out = np.full((2, n_tris), MISSING_ID, dtype=np.int64)
    fill = np.zeros(n_tris, dtype=np.int64)
    for t in range(tet_to_tri.shape[1]):
        for k in range(4):
            f = int(tet_to_tri[k, t])
            if fill[f] < 2:
                out[fill[f], f] = t
            fill[f] += 1
    return out


def pec_dof_set(tet_to_field: np.ndarray, conductor_tets, pec_tris, tri_to_edge_tab,
                edge_to_field: np.ndarray, tri_to_field: np.ndarray) -> set:
    """assembler.py:533-564."""
    pec_ids: list[int] = []
    for itet in conductor_tets:
        pec_ids.extend(int(x) for x in tet_to_field[:, itet])
    tri_ids = np.asarray(list(pec_tris), dtype=np.int64)
    if tri_ids.size:
        edge_ids = list(tri_to_edge_tab[:, tri_ids].flatten())
        for ii in edge_ids:
            pec_ids.extend(int(x) for x in edge_to_field[:, ii])
        for ii in tri_ids:
            pec_ids.extend(int(x) for x in tri_to_field[:, ii])
    return set(pec_ids)


def solve_ids(n_field: int, pec_ids: set) -> np.ndarray:
    """assembler.py:742."""
    return np.array([i for i in range(n_field) if i not in pec_ids], dtype=np.int64)


def select_block(indptr, indices, data, ids, n):
    """CSC arrays of A[:, ids][ids, :] for an n x n CSC matrix and ascending ids."""
    new = np.full(n, -1, dtype=np.int64)
    new[ids] = np.arange(len(ids))
    ip = [0]
    ix, dx = [], []
    for c in ids:
        for j in range(indptr[c], indptr[c + 1]):
            r = new[indices[j]]
            if r >= 0:
                ix.append(r)
                dx.append(data[j])
        ip.append(len(ix))
    return (np.array(ip, dtype=np.int64), np.array(ix, dtype=np.int64),
            np.array(dx, dtype=np.asarray(data).dtype))
