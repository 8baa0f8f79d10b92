"""Generate ``tri_tables.inc``: the constant table of the Nedelec-2 surface (Robin) blocks.

Run ``python gen_tri_tables.py`` in this directory; the output is committed and included
by ``robin.hip``. Exact rational arithmetic, rounded once to double.

A boundary triangle's vertices are its node ids ascending (the mesh's sorted face
tuples). Its 8 functions, in the order of the reference's 2D DOF codes
``[64]*3 + [66]*3 + [128, 129]`` (elements/dofsets.py:39 with dim 2) and local vertex
pairs (0,1), (1,2), (0,2) for the edge index (physics/microwave/assembly/robinbc.py:
256-272), are

  64:  l_i g_j - l_j g_i            128: l_0 (l_1 g_2 - l_2 g_1)
  66:  (l_i - l_j)(l_i g_j - l_j g_i) 129: l_1 (l_0 g_2 - l_2 g_0)

with g_a the surface gradient of the barycentric l_a. Writing f_i = sum_a P_ia(l) g_a,
g_a . g_b = (e_a . e_b) / (4 A^2) with e_0 = v2 - v1, e_1 = v0 - v2, e_2 = v1 - v0, and
int_T l^alpha dA = 2A alpha! / (|alpha| + 2)!, the block is

  B_ij = gamma int_T f_i . f_j dA = gamma / (2A) sum_c E_c TRI_E[8 i + j][c]

over the symmetric components c = (00, 11, 22, 01, 02, 12) of E_ab = e_a . e_b.
"""
from __future__ import annotations

import os
from fractions import Fraction
from math import factorial

SYM = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
EDGES = ((0, 1), (1, 2), (0, 2))


def _mono(*idx):
    e = [0, 0, 0]
    for i in idx:
        e[i] += 1
    return tuple(e)


def _add(poly, mono, c):
    poly[mono] = poly.get(mono, Fraction(0)) + c


def functions():
    """f_i as {a: {monomial: coefficient}} (coefficient polynomial of g_a)."""
    out = []
    for code in (64, 66):
        for (i, j) in EDGES:
            f = {i: {}, j: {}}
            if code == 64:
                _add(f[j], _mono(i), Fraction(1))
                _add(f[i], _mono(j), Fraction(-1))
            else:  # (l_i - l_j) l_i g_j - (l_i - l_j) l_j g_i
                _add(f[j], _mono(i, i), Fraction(1))
                _add(f[j], _mono(i, j), Fraction(-1))
                _add(f[i], _mono(i, j), Fraction(-1))
                _add(f[i], _mono(j, j), Fraction(1))
            out.append(f)
    f = {2: {}, 1: {}}  # 128: l0 l1 g2 - l0 l2 g1
    _add(f[2], _mono(0, 1), Fraction(1))
    _add(f[1], _mono(0, 2), Fraction(-1))
    out.append(f)
    f = {2: {}, 0: {}}  # 129: l1 l0 g2 - l1 l2 g0
    _add(f[2], _mono(0, 1), Fraction(1))
    _add(f[0], _mono(1, 2), Fraction(-1))
    out.append(f)
    return out


def _integral(m):
    """int_T l^m dA / (2A)."""
    return Fraction(factorial(m[0]) * factorial(m[1]) * factorial(m[2]), factorial(sum(m) + 2))


def _J(pa, pb):
    s = Fraction(0)
    for ma, ca in pa.items():
        for mb, cb in pb.items():
            s += ca * cb * _integral(tuple(x + y for x, y in zip(ma, mb)))
    return s


def table():
    fs = functions()
    tab = []
    for i in range(8):
        for j in range(8):
            row = []
            for a, b in SYM:
                v = _J(fs[i].get(a, {}), fs[j].get(b, {}))
                if a != b:
                    v += _J(fs[i].get(b, {}), fs[j].get(a, {}))
                row.append(v)
            tab.append(row)
    return tab


def render() -> str:
    lines = ["// Generated by gen_tri_tables.py -- do not edit.",
             "// TRI_E[8 i + j][c]: B_ij = gamma / (2A) sum_c (e_a . e_b) TRI_E[8 i + j][c],",
             "// c over (00, 11, 22, 01, 02, 12).",
             "#pragma once", "", "static constexpr double TRI_E[64][6] = {"]
    for row in table():
        lines.append("  {" + ", ".join(repr(float(v)) for v in row) + "},")
    lines.append("};")
    return "\n".join(lines) + "\n"


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "tri_tables.inc"), "w") as fh:
        fh.write(render())
