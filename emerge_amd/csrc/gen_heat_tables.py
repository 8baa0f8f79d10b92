"""Generate ``heat_tables.inc``: constant tables of the Lagrange-2 heat-conduction assembly.

Run ``python gen_heat_tables.py`` in this directory; the output is committed and included
by ``heat.hip``.

The reference (src/emerge/_emerge/physics/heatconduction/assembly/grad.py:139-338)
integrates with an 11-point Keast rule whose values are rounded to 8 digits (its weights
sum to 1 + 3e-7), so its element matrices are that rule applied to polynomials, not exact
integrals. To give the same matrices the tables here are that rule, evaluated exactly
(Fractions of the decimal values) and rounded once to double:

* the 10 functions (stored vertex order; grad.py:178 LOCAL_EDGE_MAP) are polynomials in
  l_a = s L_a, where L_a are the barycentrics and s = sign(det J): the reference divides
  its cofactors by 6|V| (grad.py:217-220), so s enters where the orientation is negative;
  N_v = l_v (2 l_v - 1), N_e = 4 l_a l_b, grad N_i = sum_a P_ia(l) grad l_a with
  P_vv = 4 l_v - 1 and P_e,a = 4 l_b, P_e,b = 4 l_a (grad.py:139-172);
* at the rule's point q, l_a = s D_aq (the point is sum_k D_kq v_k and sum_k D_kq = 1);
* HEAT_Q[s][10 i + j][4 a + b] = sum_q w_q P_ia(l_q) P_jb(l_q): K_ij = V sum_ab
  (grad L_b . C grad L_a) HEAT_Q[s][ij][ab] (grad.py:239-261: C grad N_i . grad N_j; the
  two signs of grad l cancel);
* HEAT_R[s][10 i + j] = sum_q w_q N_i(l_q) N_j(l_q): M_ij = rho_cp V HEAT_R[s][ij].

s index 0 is s = +1, index 1 is s = -1.
"""
from __future__ import annotations

import os
from fractions import Fraction

# grad.py:179-190, as decimal strings (weights, then L1..L4 of the 11 points)
_W = ["-0.078933"] + ["0.04573333"] * 4 + ["0.14933333"] * 6
_L = [
    ["0.25", "0.78571429", "0.07142857", "0.07142857", "0.07142857", "0.39940358",
     "0.39940358", "0.39940358", "0.10059642", "0.10059642", "0.10059642"],
    ["0.25", "0.07142857", "0.07142857", "0.07142857", "0.78571429", "0.10059642",
     "0.10059642", "0.39940358", "0.39940358", "0.39940358", "0.10059642"],
    ["0.25", "0.07142857", "0.07142857", "0.78571429", "0.07142857", "0.39940358",
     "0.10059642", "0.10059642", "0.39940358", "0.10059642", "0.39940358"],
    ["0.25", "0.07142857", "0.78571429", "0.07142857", "0.07142857", "0.10059642",
     "0.39940358", "0.10059642", "0.10059642", "0.39940358", "0.39940358"],
]
EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))


def rule():
    return [Fraction(w) for w in _W], [[Fraction(x) for x in row] for row in _L]


def _N(i, l):
    if i < 4:
        return l[i] * (2 * l[i] - 1)
    a, b = EDGES[i - 4]
    return 4 * l[a] * l[b]


def _P(i, a, l):
    if i < 4:
        return 4 * l[i] - 1 if a == i else Fraction(0)
    p, q = EDGES[i - 4]
    if a == p:
        return 4 * l[q]
    if a == q:
        return 4 * l[p]
    return Fraction(0)


def tables():
    W, L = rule()
    Q, R = [], []
    for s in (1, -1):
        pts = [[s * L[k][q] for k in range(4)] for q in range(11)]
        Qs = [[sum(W[q] * _P(i, a, pts[q]) * _P(j, b, pts[q]) for q in range(11))
               for a in range(4) for b in range(4)] for i in range(10) for j in range(10)]
        Rs = [sum(W[q] * _N(i, pts[q]) * _N(j, pts[q]) for q in range(11))
              for i in range(10) for j in range(10)]
        Q.append(Qs)
        R.append(Rs)
    return Q, R


def render() -> str:
    Q, R = tables()
    out = ["// Generated by gen_heat_tables.py -- do not edit.", "#pragma once", "",
           "static constexpr double HEAT_Q[2][100][16] = {"]
    for Qs in Q:
        out.append("  {")
        for row in Qs:
            out.append("    {" + ", ".join(repr(float(v)) for v in row) + "},")
        out.append("  },")
    out.append("};")
    out.append("static constexpr double HEAT_R[2][100] = {")
    for Rs in R:
        out.append("  {" + ", ".join(repr(float(v)) for v in Rs) + "},")
    out.append("};")
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "heat_tables.inc"), "w") as fh:
        fh.write(render())
