"""ORACLE -- test infrastructure, not product code.

Python side of the heat-conduction oracle: ctypes binding of ``heat_oracle.c`` (the
restatement of physics/heatconduction/assembly/grad.py) and what pins it:

* ``REFERENCE_RULE`` -- the reference's 8-digit 11-point table (grad.py:179-190);
* ``exact_rule`` -- the same Keast rule in closed form (weights -148/1875, 343/7500,
  56/375; points 11/14, 1/14 and (1 +- sqrt(5/14))/4), exact for degree 4;
* ``exact_element`` -- the Lagrange-2 stiffness and mass of one tet by exact rational
  integration of the functions' definitions (int L^a dV = 6V a! / (|a| + 3)!).
"""
from __future__ import annotations

import ctypes
from fractions import Fraction
from math import factorial

import numpy as np

from . import _load, _ptr, _c

P = ctypes.c_void_p
I64 = ctypes.c_int64

REFERENCE_RULE = np.array([
    [-0.078933, 0.04573333, 0.04573333, 0.04573333, 0.04573333,
     0.14933333, 0.14933333, 0.14933333, 0.14933333, 0.14933333, 0.14933333],
    [0.25, 0.78571429, 0.07142857, 0.07142857, 0.07142857,
     0.39940358, 0.39940358, 0.39940358, 0.10059642, 0.10059642, 0.10059642],
    [0.25, 0.07142857, 0.07142857, 0.07142857, 0.78571429,
     0.10059642, 0.10059642, 0.39940358, 0.39940358, 0.39940358, 0.10059642],
    [0.25, 0.07142857, 0.07142857, 0.78571429, 0.07142857,
     0.39940358, 0.10059642, 0.10059642, 0.39940358, 0.10059642, 0.39940358],
    [0.25, 0.07142857, 0.78571429, 0.07142857, 0.07142857,
     0.10059642, 0.39940358, 0.10059642, 0.10059642, 0.39940358, 0.39940358],
], dtype=np.float64)


def exact_rule() -> np.ndarray:
    a, b = 11 / 14, 1 / 14
    c, d = (1 + np.sqrt(5 / 14)) / 4, (1 - np.sqrt(5 / 14)) / 4
    R = REFERENCE_RULE.copy()
    R[0] = [-148 / 1875] + [343 / 7500] * 4 + [56 / 375] * 6
    sub = {0.78571429: a, 0.07142857: b, 0.39940358: c, 0.10059642: d, 0.25: 0.25}
    for r in range(1, 5):
        R[r] = [sub[v] for v in REFERENCE_RULE[r]]
    return R


def _lib():
    lib = _load()
    if not getattr(lib, "_heat_ready", False):
        lib.heat_oracle_build.restype = None
        lib.heat_oracle_build.argtypes = [P, I64, P, I64, P, P, P, P, P, P, P, P, ctypes.c_int]
        lib._heat_ready = True
    return lib


def build(nodes, tets, tet_to_field, cond=None, rho_cp=None, rule=None, rowcol=False,
          local_frame=False):
    """(K, M, rows, cols) of grad.py's builders; K needs cond (3, 3, T), M rho_cp (T,)."""
    nodes = _c(nodes, np.float64)
    tets = _c(tets, np.int64)
    t2f = _c(tet_to_field, np.int64)
    T = tets.shape[1]
    rule = _c(REFERENCE_RULE if rule is None else rule, np.float64)
    K = np.empty(100 * T) if cond is not None else None
    M = np.empty(100 * T) if rho_cp is not None else None
    cnd = None if cond is None else _c(cond, np.float64)
    rho = None if rho_cp is None else _c(rho_cp, np.float64)
    rows = np.empty(100 * T, np.int64) if rowcol else None
    cols = np.empty(100 * T, np.int64) if rowcol else None
    f = lambda a: None if a is None else _ptr(a)  # noqa: E731
    _lib().heat_oracle_build(_ptr(nodes), nodes.shape[1], _ptr(tets), T, _ptr(t2f), f(cnd),
                             f(rho), _ptr(rule), f(K), f(M), f(rows), f(cols), int(local_frame))
    return K, M, rows, cols


_EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))


def _mono(*idx):
    e = [0, 0, 0, 0]
    for i in idx:
        e[i] += 1
    return tuple(e)


def _shape(i):
    """N_i as {monomial: coefficient} in the true barycentrics."""
    if i < 4:
        return {_mono(i, i): Fraction(2), _mono(i): Fraction(-1)}
    a, b = _EDGES[i - 4]
    return {_mono(a, b): Fraction(4)}


def _dshape(i, a):
    """dN_i / dL_a as {monomial: coefficient}."""
    out = {}
    for m, c in _shape(i).items():
        if m[a]:
            mm = list(m)
            mm[a] -= 1
            out[tuple(mm)] = out.get(tuple(mm), Fraction(0)) + c * m[a]
    return out


def _int(m):  # int_T L^m dV / (6V)
    return Fraction(factorial(m[0]) * factorial(m[1]) * factorial(m[2]) * factorial(m[3]),
                    factorial(sum(m) + 3))


def _prod_int(p, q):
    return sum(cp * cq * _int(tuple(x + y for x, y in zip(mp, mq)))
               for mp, cp in p.items() for mq, cq in q.items())


def exact_element(v: np.ndarray, cond: np.ndarray, rho_cp: float):
    """Exact (K, M) of one tet: v (3, 4) vertex coordinates (any orientation)."""
    J = np.array([v[:, k] - v[:, 0] for k in (1, 2, 3)]).T
    V = abs(np.linalg.det(J)) / 6
    Ji = np.linalg.inv(J)
    g = np.vstack([-Ji.sum(axis=0), Ji])  # grad L_a, rows
    K = np.zeros((10, 10))
    M = np.zeros((10, 10))
    for i in range(10):
        for j in range(10):
            M[i, j] = float(_prod_int(_shape(i), _shape(j))) * 6 * V * rho_cp
            s = 0.0
            for a in range(4):
                for b in range(4):
                    w = float(_prod_int(_dshape(i, a), _dshape(j, b)))
                    if w:
                        s += w * (g[b] @ cond @ g[a])
            K[i, j] = s * 6 * V
    return K, M
