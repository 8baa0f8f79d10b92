"""ORACLE -- test infrastructure, not product code.

CPU restatement of the reference's Nedelec-2 curl-curl assembly
(src/emerge/_emerge/physics/microwave/assembly/curlcurl.py:222-456) and its COO->CSC
map (src/emerge/_emerge/mth/csc_cast.py:29-194), in C with OpenMP
(``ned2_oracle.c``), plus an exact rational restatement of the element integrals
(``closed_form.py``) that pins it.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``
may use this package, and only as the checker or the timed CPU baseline. The product
package (``emerge_amd``) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_DEFAULT_LIB = os.path.join(_HERE, "libned2_oracle.so")
# EMERGE_ORACLE_LIB selects another build of the same sources (the sanitizer build of
# `make -C oracle sanitize`); it is never built implicitly.
_LIB_PATH = os.environ.get("EMERGE_ORACLE_LIB", _DEFAULT_LIB)
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)
    return _DEFAULT_LIB


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        if _LIB_PATH != _DEFAULT_LIB:
            raise OSError(f"EMERGE_ORACLE_LIB={_LIB_PATH} does not exist")
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    P = ctypes.c_void_p
    i64 = ctypes.c_int64
    lib.ned2_oracle_build.restype = ctypes.c_int
    lib.ned2_oracle_build.argtypes = [P, i64, P, i64, P, i64, P, i64, P, P, P, P, P, P, i64,
                                      P, ctypes.c_int, P, P, P, P, ctypes.c_int,
                                      ctypes.c_int]
    lib.ned2_oracle_csc_pattern.restype = i64
    lib.ned2_oracle_csc_pattern.argtypes = [P, P, i64, i64, P, P, P]
    lib.ned2_oracle_csc_fill.restype = None
    lib.ned2_oracle_csc_fill.argtypes = [P, i64, P, P, P, P, P]
    lib.ned2_oracle_csc_scatter.restype = None
    lib.ned2_oracle_csc_scatter.argtypes = [P, P, P, P, i64, P, i64]
    lib.ned2_oracle_keast.restype = None
    lib.ned2_oracle_keast.argtypes = [P]
    _lib = lib
    return lib


def keast_table() -> np.ndarray:
    """The oracle's 11-point rule (5, 11): weights, then barycentric coordinates
    (curlcurl.py:277-283)."""
    out = np.empty((5, 11), np.float64)
    _load().ned2_oracle_keast(_ptr(out))
    return out


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class OracleError(RuntimeError):
    pass


def matrix_builder(nodes, tets, tris, edges, tet_to_field, tet_to_edge, tet_to_tri,
                   ur, er, tetids, dofcodes, nthreads: int | None = None,
                   local_frame: bool = False):
    """Restatement of ``_matrix_builder`` (curlcurl.py:387-456), same argument order
    (``ur`` before ``er``). Returns (dataE, dataB, rows, cols).

    ``local_frame=False`` repeats the reference's arithmetic: the steps it compiles
    without fastmath (tet_coefficients, matinv, the basis functions) in its own operand
    order, in absolute coordinates (``ned2_strict.c``). ``local_frame=True`` evaluates
    each tet in coordinates relative to its stored vertex 0 (same algorithm; avoids the
    cancellation of the absolute-coordinate cofactor sums on thin tets, which puts the
    reference's values 1.1e-11 to 3.6e-7 relative Frobenius from the exact integrals on
    seeded Delaunay meshes of 21k to 2.8M tets)."""
    lib = _load()
    nodes = _c(nodes, np.float64)
    tets = _c(tets, np.int64)
    tris = _c(tris, np.int64)
    edges = _c(edges, np.int64)
    t2f = _c(tet_to_field, np.int64)
    t2e = _c(tet_to_edge, np.int64)
    t2t = _c(tet_to_tri, np.int64)
    ur = _c(ur, np.complex128)
    er = _c(er, np.complex128)
    tetids = _c(tetids, np.int64)
    dofcodes = _c(dofcodes, np.int64)
    T = tets.shape[1]
    ndof = dofcodes.shape[0]
    Ta = tetids.shape[0]
    n = Ta * ndof * ndof
    dataE = np.empty(n, np.complex128)
    dataB = np.empty(n, np.complex128)
    rows = np.empty(n, np.int64)
    cols = np.empty(n, np.int64)
    nt = nthreads if nthreads is not None else (os.cpu_count() or 1)
    st = lib.ned2_oracle_build(_ptr(nodes), nodes.shape[1], _ptr(tets), T, _ptr(tris),
                               tris.shape[1], _ptr(edges), edges.shape[1], _ptr(t2f),
                               _ptr(t2e), _ptr(t2t), _ptr(ur), _ptr(er), _ptr(tetids), Ta,
                               _ptr(dofcodes), ndof, _ptr(dataE), _ptr(dataB), _ptr(rows),
                               _ptr(cols), int(nt), int(bool(local_frame)))
    if st != 0:
        raise OracleError(f"oracle builder status {st}")
    return dataE, dataB, rows, cols


def tet_mass_stiffness_matrices(field, er, ur, conductor_tets, nthreads=None,
                                local_frame=False):
    """Restatement of ``tet_mass_stiffness_matrices`` (curlcurl.py:222-270) without the
    CSC map: returns (dataE, dataB, rows, cols)."""
    T = field.mesh.tets.shape[1]
    tet_assy = np.arange(T)
    conductor_tets = np.asarray(conductor_tets)
    if conductor_tets.shape[0] > 0:
        tet_assy = np.delete(tet_assy, conductor_tets.astype(np.int64))
    m = field.mesh
    return matrix_builder(m.nodes, m.tets, m.tris, m.edges, field.get_tet_to_field(),
                          m.tet_to_edge, m.tet_to_tri, ur, er, tet_assy, field.dofcodes3d,
                          nthreads=nthreads, local_frame=local_frame)


class CSCMap:
    """Restatement of ``CSCMapping`` (csc_cast.py:29-63)."""

    def __init__(self, rows, cols, N):
        lib = _load()
        rows = _c(rows, np.int64)
        cols = _c(cols, np.int64)
        nnz_coo = rows.shape[0]
        self.N = int(N)
        self.col_offsets = np.empty(N + 1, np.int64)
        self.perm = np.empty(nnz_coo, np.int64)
        self.indptr = np.empty(N + 1, np.int64)
        nnz = lib.ned2_oracle_csc_pattern(_ptr(rows), _ptr(cols), nnz_coo, N,
                                          _ptr(self.col_offsets), _ptr(self.perm),
                                          _ptr(self.indptr))
        self.indices = np.empty(nnz, np.int64)
        self.target_csc = np.empty(nnz_coo, np.int64)
        lib.ned2_oracle_csc_fill(_ptr(rows), N, _ptr(self.col_offsets), _ptr(self.perm),
                                 _ptr(self.indptr), _ptr(self.indices), _ptr(self.target_csc))
        self.nnz = int(nnz)
        self.nnz_coo = nnz_coo

    def to_csc_data(self, data):
        lib = _load()
        data = _c(data, np.complex128)
        out = np.empty(self.nnz, np.complex128)
        lib.ned2_oracle_csc_scatter(_ptr(data), _ptr(self.perm), _ptr(self.target_csc),
                                    _ptr(self.col_offsets), self.N, _ptr(out), self.nnz)
        return out

    def to_csc(self, data):
        from scipy.sparse import csc_matrix
        return csc_matrix((self.to_csc_data(data), self.indices, self.indptr),
                          shape=(self.N, self.N))
