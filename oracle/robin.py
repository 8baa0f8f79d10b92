"""ORACLE -- test infrastructure, not product code.

Python side of the surface (Robin / port) oracle: ctypes bindings of ``tri_oracle.c``
(the step-by-step restatement of robinbc.py) and an exact rational restatement of the
8x8 face blocks that pins it:

* ``bc_entries`` -- compute_bc_entries (robinbc.py:385-402) through the C restatement;
* ``force_entries`` -- compute_force_entries[_scat] (robinbc.py:278-313, 517-569);
* ``exact_block`` -- int_T f_i . f_j dA from the functions' definitions
  (elements/dofsets.py:66-83) with int_T l^a dA = 2A a! / (|a| + 2)!, evaluated with
  Fractions for the integrals and the face's edge vectors for the gradients.
"""
from __future__ import annotations

import ctypes
from fractions import Fraction
from math import factorial

import numpy as np

from . import _load, _ptr, _c

P = ctypes.c_void_p
I64 = ctypes.c_int64


def _lib():
    lib = _load()
    if not getattr(lib, "_tri_ready", False):
        lib.tri_oracle_bc_entries.restype = None
        lib.tri_oracle_bc_entries.argtypes = [P, I64, P, I64, P, I64, ctypes.c_double,
                                              ctypes.c_double, P]
        lib.tri_oracle_force_entries.restype = None
        lib.tri_oracle_force_entries.argtypes = [P, I64, P, I64, P, P, I64, P, P, P, P]
        lib.tri_oracle_dpts.restype = None
        lib.tri_oracle_dpts.argtypes = [P]
        lib._tri_ready = True
    return lib


def dpts() -> np.ndarray:
    out = np.empty((4, 6), np.float64)
    _lib().tri_oracle_dpts(_ptr(out))
    return out


def bc_entries(nodes, tris, surf, gamma, Bmat=None):
    """Bmat (n_tris * 64) + the blocks of the listed faces (robinbc.py:385-402)."""
    nodes = _c(nodes, np.float64)
    tris = _c(tris, np.int64)
    surf = _c(surf, np.int64).reshape(-1)
    F = tris.shape[1]
    B = np.zeros(F * 64, np.complex128) if Bmat is None else _c(Bmat, np.complex128).copy()
    g = complex(gamma)
    _lib().tri_oracle_bc_entries(_ptr(nodes), nodes.shape[1], _ptr(tris), F, _ptr(surf),
                                 surf.shape[0], g.real, g.imag, _ptr(B))
    return B


def force_entries(nodes, tris, tri_to_field, n_field, surf, U, U_curl=None, normals=None):
    """The excitation vector (robinbc.py:278-313, or :517-569 with U_curl/normals)."""
    nodes = _c(nodes, np.float64)
    tris = _c(tris, np.int64)
    t2f = _c(tri_to_field, np.int64)
    surf = _c(surf, np.int64).reshape(-1)
    U = _c(U, np.complex128)
    Uc = None if U_curl is None else _c(U_curl, np.complex128)
    nh = None if normals is None else _c(normals, np.float64)
    out = np.zeros(n_field, np.complex128)
    _lib().tri_oracle_force_entries(_ptr(nodes), nodes.shape[1], _ptr(tris), tris.shape[1],
                                    _ptr(t2f), _ptr(surf), surf.shape[0], _ptr(U),
                                    None if Uc is None else _ptr(Uc),
                                    None if nh is None else _ptr(nh), _ptr(out))
    return out


# ---- exact restatement ---------------------------------------------------------------

_EDGES = ((0, 1), (1, 2), (0, 2))


def _mono(*idx):
    e = [0, 0, 0]
    for i in idx:
        e[i] += 1
    return tuple(e)


def _functions():
    """The 8 functions as {gradient index a: {monomial: coefficient}}."""
    fs = []
    for code in (64, 66):
        for i, j in _EDGES:
            f = {i: {}, j: {}}
            if code == 64:  # l_i g_j - l_j g_i
                f[j][_mono(i)] = Fraction(1)
                f[i][_mono(j)] = Fraction(-1)
            else:  # (l_i - l_j)(l_i g_j - l_j g_i)
                f[j][_mono(i, i)] = Fraction(1)
                f[j][_mono(i, j)] = Fraction(-1)
                f[i][_mono(i, j)] = Fraction(-1)
                f[i][_mono(j, j)] = Fraction(1)
            fs.append(f)
    fs.append({2: {_mono(0, 1): Fraction(1)}, 1: {_mono(0, 2): Fraction(-1)}})  # 128
    fs.append({2: {_mono(0, 1): Fraction(1)}, 0: {_mono(1, 2): Fraction(-1)}})  # 129
    return fs


def _int(m):  # int_T l^m dA / (2A)
    return Fraction(factorial(m[0]) * factorial(m[1]) * factorial(m[2]), factorial(sum(m) + 2))


def exact_block(v: np.ndarray) -> np.ndarray:
    """int_T f_i . f_j dA (8x8) on the face with vertex coordinates v (3 x 3, columns in
    ascending node-id order). Gradients from the edge vectors: g_a . g_b =
    (e_a . e_b) / (4 A^2), e_0 = v2 - v1, e_1 = v0 - v2, e_2 = v1 - v0."""
    fs = _functions()
    p = [v[:, k] for k in range(3)]
    e = [p[2] - p[1], p[0] - p[2], p[1] - p[0]]
    A = 0.5 * np.linalg.norm(np.cross(p[1] - p[0], p[2] - p[0]))
    G = np.array([[e[a] @ e[b] for b in range(3)] for a in range(3)]) / (4 * A * A)
    out = np.zeros((8, 8))
    for i in range(8):
        for j in range(8):
            s = 0.0
            for a, pa in fs[i].items():
                for b, pb in fs[j].items():
                    J = sum(ca * cb * _int(tuple(x + y for x, y in zip(ma, mb)))
                            for ma, ca in pa.items() for mb, cb in pb.items())
                    s += float(J) * G[a, b]
            out[i, j] = 2 * A * s
    return out
