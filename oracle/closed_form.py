"""ORACLE -- test infrastructure, not product code.

Exact element matrices of the reference's Nedelec-2 tetrahedral curl-curl forms, in
rational arithmetic. Used to pin the C restatement (``ned2_oracle.c``): the reference's
11-point Keast rule (physics/microwave/assembly/curlcurl.py:277-283) is exact for
polynomials of degree <= 4 and every integrand here has degree <= 4, so the
reference's element matrices equal these exact integrals up to rounding.

What is restated, and from where (paths under src/emerge/_emerge/):

* basis functions from their definitions in elements/dofsets.py:66-83, codes 64/66
  (edge functions 0 and 2) and 128/129 (face functions 0 and 1);
* the vertex order inside each edge/face function is ascending global node id
  (edges/tris are stored as sorted tuples, mesh3d.py:714-732, mapped back to local
  positions by mth/optimized.py:262-296) -- here it is derived from the ids directly,
  not from the tables, so it checks the tables too;
* the |V| convention of mth/optimized.py:142-167 with curlcurl.py:306-311: the
  reference's barycentric functions are ``s * lambda`` with ``s = sign(det J)``;
* K_ij = int (mu^-1 curl f_i) . curl f_j, M_ij = int (eps f_i) . f_j for i <= j in slot
  order, mirrored (curlcurl.py:358-376).

Integration uses int lambda^alpha dV = alpha! 3! |V| / (|alpha| + 3)!.
"""
from __future__ import annotations

from fractions import Fraction
from math import factorial

EDGE_SLOTS = ((0, 1), (0, 2), (0, 3), (1, 2), (3, 1), (2, 3))
FACE_SLOTS = ((0, 1, 2), (0, 2, 3), (0, 3, 1), (1, 2, 3))
CODES = [64] * 6 + [66] * 6 + [128] * 4 + [129] * 4


def _det3(m):
    return (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1])
            - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
            + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))


def _inv3(m):
    d = _det3(m)
    c = [[None] * 3 for _ in range(3)]
    for i in range(3):
        for j in range(3):
            minor = [[m[r][s] for s in range(3) if s != j] for r in range(3) if r != i]
            c[i][j] = (-1) ** (i + j) * (minor[0][0] * minor[1][1] - minor[0][1] * minor[1][0])
    return [[c[j][i] / d for j in range(3)] for i in range(3)]


def _cross(u, v):
    return [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]]


def _mono(*idx):
    e = [0, 0, 0, 0]
    for i in idx:
        e[i] += 1
    return tuple(e)


class VPoly:
    """Vector polynomial in barycentric variables: {exponent (a0..a3): [vx, vy, vz]}."""

    def __init__(self, terms=None):
        self.t = {}
        for m, v in (terms or {}).items():
            self.add(m, v)

    def add(self, m, v, scale=1):
        cur = self.t.setdefault(m, [0, 0, 0])
        for k in range(3):
            cur[k] += scale * v[k]

    def curl(self, grads):
        """curl(p(lambda) v) = sum_n dp/dlambda_n grad(lambda_n) x v."""
        out = VPoly()
        for m, v in self.t.items():
            for n in range(4):
                if m[n] == 0:
                    continue
                dm = list(m)
                dm[n] -= 1
                out.add(tuple(dm), _cross(grads[n], v), scale=m[n])
        return out


def basis(code, i, j, k, g):
    """Basis function ``code`` on local vertices (i, j, k) with gradients ``g``.

    ``g`` are gradients of the functions the basis is built from (here ``s*lambda``),
    and the polynomial part is in ``lambda``; the caller multiplies by ``s`` once per
    barycentric factor of the polynomial part.
    Returns (VPoly f, total degree counting the gradient factor).
    """
    f = VPoly()
    if code == 64:   # L_i grad L_j - L_j grad L_i
        f.add(_mono(i), g[j])
        f.add(_mono(j), g[i], -1)
        return f, 2
    if code == 66:   # (L_i - L_j)(L_i grad L_j - L_j grad L_i)
        f.add(_mono(i, i), g[j])
        f.add(_mono(i, j), g[i], -1)
        f.add(_mono(j, i), g[j], -1)
        f.add(_mono(j, j), g[i])
        return f, 3
    if code == 128:  # L_i (L_j grad L_k - L_k grad L_j)
        f.add(_mono(i, j), g[k])
        f.add(_mono(i, k), g[j], -1)
        return f, 3
    if code == 129:  # L_j (L_i grad L_k - L_k grad L_i)
        f.add(_mono(j, i), g[k])
        f.add(_mono(j, k), g[i], -1)
        return f, 3
    raise ValueError(f"unsupported DOF code {code}")


def _int_mono(m, absV):
    num = 1
    for a in m:
        num *= factorial(a)
    return Fraction(num * 6, factorial(sum(m) + 3)) * absV


def _bilinear(fa: VPoly, fb: VPoly, mat, absV):
    """int (mat fa) . fb dV."""
    acc = Fraction(0)
    for ma, va in fa.t.items():
        mva = [sum(mat[r][c] * va[c] for c in range(3)) for r in range(3)]
        for mb, vb in fb.t.items():
            dot = mva[0] * vb[0] + mva[1] * vb[1] + mva[2] * vb[2]
            if dot == 0:
                continue
            m = tuple(ma[n] + mb[n] for n in range(4))
            acc += dot * _int_mono(m, absV)
    return acc


def element_matrices(verts, gids, eps, mu, mu_is_inverse=False):
    """Exact (K, M) 20x20 of one tet, reference slot order.

    verts: 4 vertex coordinates (rationals) in stored order; gids: their global ids;
    eps, mu: 3x3 rational tensors (complex tensors: call once per real part, the forms
    are linear in eps and mu^-1). The reference uses K with mu^-1
    (mth/optimized.py:315 matinv) and M with eps; ``mu_is_inverse`` passes mu^-1.
    """
    v = [[Fraction(c) for c in p] for p in verts]
    J = [[v[k][r] - v[0][r] for k in (1, 2, 3)] for r in range(3)]
    detJ = _det3(J)
    if detJ == 0:
        raise ZeroDivisionError("degenerate tet")
    s = 1 if detJ > 0 else -1
    absV = abs(detJ) / 6
    Jinv = _inv3(J)
    # true barycentric gradients: grad lambda_k = row k-1 of J^-1 (k=1..3)
    gl = [None] * 4
    for k in (1, 2, 3):
        gl[k] = [Jinv[k - 1][c] for c in range(3)]
    gl[0] = [-(gl[1][c] + gl[2][c] + gl[3][c]) for c in range(3)]
    g = [[s * x for x in gl[n]] for n in range(4)]   # reference: s * lambda
    muq = [[Fraction(x) for x in row] for row in mu]
    muinv = muq if mu_is_inverse else _inv3(muq)
    epsq = [[Fraction(x) for x in row] for row in eps]
    fs, cs = [], []
    for d, code in enumerate(CODES):
        if d < 12:
            a, b = EDGE_SLOTS[d % 6]
            i, j = (a, b) if gids[a] < gids[b] else (b, a)
            k = 0
        else:
            tri = sorted(FACE_SLOTS[(d - 12) % 4], key=lambda p: gids[p])
            i, j, k = tri
        f, deg = basis(code, i, j, k, g)
        # the polynomial part is in s*lambda: one factor s per barycentric factor
        npoly = deg - 1
        sp = s ** npoly
        f = VPoly({m: [sp * x for x in vv] for m, vv in f.t.items()})
        fs.append(f)
        # the polynomial variables are lambda, so the chain rule uses grad lambda
        cs.append(f.curl(gl))
    K = [[Fraction(0)] * 20 for _ in range(20)]
    M = [[Fraction(0)] * 20 for _ in range(20)]
    for a in range(20):
        for b in range(a, 20):
            K[a][b] = K[b][a] = _bilinear(cs[a], cs[b], muinv, absV)
            M[a][b] = M[b][a] = _bilinear(fs[a], fs[b], epsq, absV)
    return K, M


def assemble_dense(nodes, tets, tet_to_field, n_field, eps=None, mu=None):
    """Exact global K, M (dicts keyed by (row, col)) summed over tets; for tiny meshes."""
    I3 = [[1, 0, 0], [0, 1, 0], [0, 0, 1]]
    eps = eps or I3
    mu = mu or I3
    Kg, Mg = {}, {}
    T = len(tets[0])
    for t in range(T):
        vid = [int(tets[v][t]) for v in range(4)]
        verts = [[nodes[c][vid[v]] for c in range(3)] for v in range(4)]
        K, M = element_matrices(verts, vid, eps, mu)
        dofs = [int(tet_to_field[d][t]) for d in range(20)]
        for a in range(20):
            for b in range(20):
                key = (dofs[a], dofs[b])
                Kg[key] = Kg.get(key, 0) + K[a][b]
                Mg[key] = Mg.get(key, 0) + M[a][b]
    return Kg, Mg
