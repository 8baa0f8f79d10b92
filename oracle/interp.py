"""ORACLE -- test infrastructure, not product code.

ctypes bindings of the C restatements of the reference's field interpolation
(``ned2_oracle_interp``: compiled/base/interp.py:286-387 ned2_tet_interp, :410-519
ned2_tet_interp_curl) and point location (``ned2_oracle_locate``: compiled/base/
index_inter.py:134-246, brute force), plus the reference's surface integral for the
S-parameters (mth/integrals.py:31-42 and :64-104, sparam.py:26-97) written out in NumPy.

Pins (tests/test_interp_oracle.py): interpolating the Whitney coefficients of a constant
field E0 and of a rotation field E0 + b x r reproduces the field and its curl b exactly;
the interpolant equals the closed-form basis sum; located tets contain their points.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _load, _ptr, _c

P = ctypes.c_void_p
I64 = ctypes.c_int64


def _lib():
    lib = _load()
    if not getattr(lib, "_interp_ready", False):
        lib.ned2_oracle_interp.restype = ctypes.c_int
        lib.ned2_oracle_interp.argtypes = [P, I64, P, I64, P, I64, P, I64, P, P, P, P,
                                           ctypes.c_int, P, I64, P, P, ctypes.c_int, P,
                                           ctypes.c_int, P]
        lib.ned2_oracle_locate.restype = None
        lib.ned2_oracle_locate.argtypes = [P, I64, P, I64, P, I64, P, I64, P]
        lib._interp_ready = True
    return lib


def interp(field, coords, sol, tet_of, curl=False, c_tet=None, local_frame=False):
    """(3, n) complex: the solution's field (or c_tet * curl) at coords (3, n)."""
    m = field.mesh
    nodes = _c(m.nodes, np.float64)
    tets = _c(m.tets, np.int64)
    tris = _c(m.tris, np.int64)
    edges = _c(m.edges, np.int64)
    t2f = _c(field.tet_to_field, np.int64)
    t2e = _c(m.tet_to_edge, np.int64)
    t2t = _c(m.tet_to_tri, np.int64)
    codes = _c(field.dofcodes3d, np.int64)
    coords = _c(coords, np.float64)
    n = coords.shape[1]
    tet_of = _c(tet_of, np.int64)
    sol = _c(sol, np.complex128)
    ct = None if c_tet is None else _c(c_tet, np.complex128)
    out = np.empty((3, n), np.complex128)
    st = _lib().ned2_oracle_interp(_ptr(nodes), nodes.shape[1], _ptr(tets), tets.shape[1],
                                   _ptr(tris), tris.shape[1], _ptr(edges), edges.shape[1],
                                   _ptr(t2f), _ptr(t2e), _ptr(t2t), _ptr(codes), codes.shape[0],
                                   _ptr(coords), n, _ptr(tet_of), _ptr(sol), int(curl),
                                   None if ct is None else _ptr(ct), int(local_frame), _ptr(out))
    if st != 0:
        raise RuntimeError(f"oracle interp status {st}")
    out[:, tet_of < 0] = np.nan
    return out


def locate(mesh, coords, tetids=None):
    nodes = _c(mesh.nodes, np.float64)
    tets = _c(mesh.tets, np.int64)
    coords = _c(coords, np.float64)
    ids = None if tetids is None else _c(tetids, np.int64)
    out = np.empty(coords.shape[1], np.int64)
    _lib().ned2_oracle_locate(_ptr(nodes), nodes.shape[1], _ptr(tets), tets.shape[1],
                              None if ids is None else _ptr(ids),
                              tets.shape[1] if ids is None else ids.shape[0],
                              _ptr(coords), coords.shape[1], _ptr(out))
    return out


def whitney_coefficients(field, E0, b=None):
    """DOF vector of E0 + b x r: the line integrals along every edge (low -> high global
    id) on the code-64 DOFs, zero elsewhere (Whitney functions span a + b x r)."""
    m = field.mesh
    p0 = m.nodes[:, m.edges[0]]
    p1 = m.nodes[:, m.edges[1]]
    t = p1 - p0
    mid = 0.5 * (p0 + p1)
    f = np.broadcast_to(np.asarray(E0, np.complex128)[:, None], t.shape).copy()
    if b is not None:
        f = f + np.cross(np.asarray(b)[None, :], mid.T).T
    sol = np.zeros(field.n_field, np.complex128)
    sol[:m.n_edges] = np.sum(f * t, axis=0)
    return sol
