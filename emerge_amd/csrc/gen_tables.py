"""Generate ``ned2_tables.inc``: the constant tables of the Nedelec-2 assembly kernels.

Run ``python gen_tables.py`` in this directory; the output is committed and included by
the HIP sources. Everything is computed in exact rational arithmetic, then rounded
once to the nearest double.

Canonical element. A tet whose four vertices are ordered by ascending global node id
is "canonical": every edge/face function then has its vertices in local order (the
reference orients each function by ascending global id). On the unit tet
xi-space, lambda_0 = 1 - xi_1 - xi_2 - xi_3, lambda_k = xi_k, and the 20 functions are,
in slot order (the reference's slot order, src/emerge/_emerge/mesh3d.py:760-761, with
code order [64]*6 + [66]*6 + [128]*4 + [129]*4, elements/dofsets.py:39):

  edges (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)      faces (0,1,2) (0,2,3) (0,1,3) (1,2,3)
  64:  l_i g_j - l_j g_i                          128: l_i (l_j g_k - l_k g_j)
  66:  (l_i - l_j)(l_i g_j - l_j g_i)             129: l_j (l_i g_k - l_k g_i)

Tables written:

* ``TSYM_K[210][6]``, ``TSYM_M[210][6]``: for the packed upper-triangle pair (a <= b)
  and symmetric component pq in (00, 11, 22, 01, 02, 12): T[a,b,p,q] (+ T[a,b,q,p]
  when p != q), where T_M[a,b,p,q] = int f_a,p f_b,q dxi and T_K the same with the
  curls.
* ``TANTI_K[210][3]``, ``TANTI_M[210][3]``: T[a,b,p,q] - T[a,b,q,p] for pq in (01, 02, 12).
* ``PAIR_A[210]``, ``PAIR_B[210]``: the pair of each packed index.
* ``SLOT_OF[64][20]``: for a vertex-rank code ``r0*16 + r1*4 + r2`` (r_m = rank of the
  global id of stored vertex m), the reference slot that holds canonical function c.
* ``PERM_PARITY[64]``: +1 / -1 parity of the rank permutation (0 for invalid codes).
"""
from __future__ import annotations

import itertools
import os
from fractions import Fraction
from math import factorial

CAN_EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))
CAN_FACES = ((0, 1, 2), (0, 2, 3), (0, 1, 3), (1, 2, 3))
# stored-vertex positions of the reference's slots
REF_EDGE_SLOTS = ((0, 1), (0, 2), (0, 3), (1, 2), (3, 1), (2, 3))
REF_FACE_SLOTS = ((0, 1, 2), (0, 2, 3), (0, 3, 1), (1, 2, 3))
SYM = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
ANTI = ((0, 1), (0, 2), (1, 2))

GRAD = ((-1, -1, -1), (1, 0, 0), (0, 1, 0), (0, 0, 1))


def _mono(*idx):
    e = [0, 0, 0, 0]
    for i in idx:
        e[i] += 1
    return tuple(e)


def _cross(u, v):
    return (u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0])


def _func(code, i, j, k):
    """Vector polynomial {monomial: vec} of a canonical function."""
    out = {}

    def add(m, v, s=1):
        cur = out.setdefault(m, [0, 0, 0])
        for c in range(3):
            cur[c] += s * v[c]

    g = GRAD
    if code == 64:
        add(_mono(i), g[j]); add(_mono(j), g[i], -1)
    elif code == 66:
        add(_mono(i, i), g[j]); add(_mono(i, j), g[i], -1)
        add(_mono(j, i), g[j], -1); add(_mono(j, j), g[i])
    elif code == 128:
        add(_mono(i, j), g[k]); add(_mono(i, k), g[j], -1)
    elif code == 129:
        add(_mono(j, i), g[k]); add(_mono(j, k), g[i], -1)
    return out


def _curl(poly):
    out = {}
    for m, v in poly.items():
        for n in range(4):
            if m[n] == 0:
                continue
            dm = list(m)
            dm[n] -= 1
            w = _cross(GRAD[n], v)
            cur = out.setdefault(tuple(dm), [0, 0, 0])
            for c in range(3):
                cur[c] += m[n] * w[c]
    return out


def _integral(m):
    # int over the unit tet (volume 1/6) of lambda^m
    num = 1
    for a in m:
        num *= factorial(a)
    return Fraction(num, factorial(sum(m) + 3))


def _tensor(fa, fb):
    """T[p][q] = int fa_p fb_q."""
    T = [[Fraction(0)] * 3 for _ in range(3)]
    for ma, va in fa.items():
        for mb, vb in fb.items():
            w = _integral(tuple(ma[n] + mb[n] for n in range(4)))
            for p in range(3):
                if va[p] == 0:
                    continue
                for q in range(3):
                    if vb[q]:
                        T[p][q] += va[p] * vb[q] * w
    return T


def canonical_functions():
    funcs = []
    for code in (64, 66):
        for (i, j) in CAN_EDGES:
            funcs.append(_func(code, i, j, 0))
    for code in (128, 129):
        for (i, j, k) in CAN_FACES:
            funcs.append(_func(code, i, j, k))
    return funcs


def pairs():
    return [(a, b) for a in range(20) for b in range(a, 20)]


def build_tables():
    fs = canonical_functions()
    cs = [_curl(f) for f in fs]
    out = {"TSYM_K": [], "TSYM_M": [], "TANTI_K": [], "TANTI_M": []}
    for a, b in pairs():
        for name, src in (("K", cs), ("M", fs)):
            T = _tensor(src[a], src[b])
            out["TSYM_" + name].append([T[p][q] + (T[q][p] if p != q else 0) for p, q in SYM])
            out["TANTI_" + name].append([T[p][q] - T[q][p] for p, q in ANTI])
    return out


def slot_tables():
    """SLOT_OF[code][c] and parity for each rank code r0*16 + r1*4 + r2."""
    slot_of = [[0] * 20 for _ in range(64)]
    parity = [0] * 64
    can_edge_index = {e: n for n, e in enumerate(CAN_EDGES)}
    can_face_index = {f: n for n, f in enumerate(CAN_FACES)}
    for ranks in itertools.permutations(range(4)):
        code = ranks[0] * 16 + ranks[1] * 4 + ranks[2]
        inv = sum(1 for x in range(4) for y in range(x + 1, 4) if ranks[x] > ranks[y])
        parity[code] = -1 if inv % 2 else 1
        ref_to_can = []
        for (m1, m2) in REF_EDGE_SLOTS:
            ref_to_can.append(can_edge_index[tuple(sorted((ranks[m1], ranks[m2])))])
        face_map = [can_face_index[tuple(sorted(ranks[m] for m in f))] for f in REF_FACE_SLOTS]
        full = ([c for c in ref_to_can] + [6 + c for c in ref_to_can]
                + [12 + c for c in face_map] + [16 + c for c in face_map])
        for slot, can in enumerate(full):
            slot_of[code][can] = slot
    return slot_of, parity


def _hex(x: Fraction) -> str:
    return float(x).hex()


def render() -> str:
    t = build_tables()
    slot_of, parity = slot_tables()
    pr = pairs()
    lines = ["// Generated by gen_tables.py -- do not edit. Exact rationals rounded to double.",
             "// Canonical Nedelec-2 (codes 64/66/128/129) reference-element tensors.",
             "#pragma once", "#define NED2_NPAIR 210", ""]
    for name in ("TSYM_K", "TSYM_M"):
        lines.append(f"static constexpr double {name}[210][6] = {{")
        for row in t[name]:
            lines.append("  {" + ", ".join(_hex(v) for v in row) + "},")
        lines.append("};")
    for name in ("TANTI_K", "TANTI_M"):
        lines.append(f"static constexpr double {name}[210][3] = {{")
        for row in t[name]:
            lines.append("  {" + ", ".join(_hex(v) for v in row) + "},")
        lines.append("};")
    lines.append("static constexpr unsigned char PAIR_A[210] = {" + ", ".join(str(a) for a, _ in pr) + "};")
    lines.append("static constexpr unsigned char PAIR_B[210] = {" + ", ".join(str(b) for _, b in pr) + "};")
    lines.append("static constexpr unsigned char SLOT_OF[64][20] = {")
    for row in slot_of:
        lines.append("  {" + ", ".join(str(v) for v in row) + "},")
    lines.append("};")
    lines.append("static constexpr signed char PERM_PARITY[64] = {" + ", ".join(str(p) for p in parity) + "};")
    # CANON[code][slot]: canonical function held by reference slot (inverse of SLOT_OF)
    canon = [[0] * 20 for _ in range(64)]
    for code in range(64):
        if parity[code] == 0:
            continue
        for c, slot in enumerate(slot_of[code]):
            canon[code][slot] = c
    lines.append("static constexpr unsigned char CANON[64][20] = {")
    for row in canon:
        lines.append("  {" + ", ".join(str(v) for v in row) + "},")
    lines.append("};")
    # PAIR_INDEX[a][b]: packed index of the unordered pair {a, b}
    index = {ab: n for n, ab in enumerate(pr)}
    lines.append("static constexpr unsigned char PAIR_INDEX[20][20] = {")
    for a in range(20):
        lines.append("  {" + ", ".join(str(index[(min(a, b), max(a, b))]) for b in range(20)) + "},")
    lines.append("};")
    return "\n".join(lines) + "\n"


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "ned2_tables.inc"), "w") as fh:
        fh.write(render())
