"""Write the golden fixtures of the Nedelec-2 element assembly.

The expected outputs are exact rational integrals (oracle/closed_form.py) rounded once
to float64; the reference itself ships no golden vectors. Run from the repository root:

    python tests/golden/make_fixtures.py

Fixtures:
* single_tet_24orders.npz -- one tet with integer vertex coordinates stored in all 24
  vertex orders (both orientations), complex anisotropic diagonal eps_r and mu_r; the
  exact 20x20 K (mu_r^-1) and M (eps_r) blocks in reference slot order.
* hex6.npz -- one unit hex split into 6 tets (hex_box_mesh), eps_r = mu_r = I; exact
  COO values of every tet's K and M block (reference layout, 400 entries per tet).
"""
from __future__ import annotations

import itertools
import os
import sys
from fractions import Fraction

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import closed_form as cf  # noqa: E402
from emerge_amd.mesh import hex_box_mesh  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

BASE = [[0, 0, 0], [3, 1, 0], [1, 4, 1], [0, 1, 5]]
EPS = np.diag([2 - 0.1j, 3 - 0.05j, 5 + 0j])
MU = np.diag([1 + 0j, 2 - 0.2j, 4 + 0j])


def _fr(m):
    return [[Fraction(float(x)) for x in row] for row in m]


def exact_complex(verts, gids, eps, muinv):
    K = np.zeros((20, 20), complex)
    M = np.zeros((20, 20), complex)
    I3 = np.eye(3).tolist()
    for unit, part in ((1.0, np.real), (1j, np.imag)):
        e, mi = part(eps), part(muinv)
        Kq, Mq = cf.element_matrices(verts, gids, _fr(e) if np.any(e) else I3,
                                     _fr(mi) if np.any(mi) else I3, mu_is_inverse=True)
        if np.any(mi):
            K += unit * np.array([[float(x) for x in r] for r in Kq])
        if np.any(e):
            M += unit * np.array([[float(x) for x in r] for r in Mq])
    return K, M


def main():
    orders = np.array(list(itertools.permutations(range(4))), np.int64)
    muinv = np.diag(1.0 / np.diag(MU))
    Ks, Ms = [], []
    for o in orders:
        verts = [[Fraction(c) for c in BASE[i]] for i in o]
        K, M = exact_complex(verts, list(o), EPS, muinv)
        Ks.append(K)
        Ms.append(M)
    np.savez_compressed(os.path.join(HERE, "single_tet_24orders.npz"), base=np.array(BASE, float),
                        orders=orders, eps=EPS, mu=MU, K=np.array(Ks), M=np.array(Ms))
    mesh = hex_box_mesh(1.0, 1.0, 1.0, 1, 1, 1)
    I3 = np.eye(3).tolist()
    Kb, Mb = [], []
    for t in range(mesh.n_tets):
        vid = [int(v) for v in mesh.tets[:, t]]
        verts = [[Fraction(float(mesh.nodes[c, v])) for c in range(3)] for v in vid]
        K, M = cf.element_matrices(verts, vid, I3, I3)
        Kb.append(np.array([[float(x) for x in r] for r in K]).ravel())
        Mb.append(np.array([[float(x) for x in r] for r in M]).ravel())
    np.savez_compressed(os.path.join(HERE, "hex6.npz"), nodes=mesh.nodes, tets=mesh.tets,
                        K=np.array(Kb), M=np.array(Mb))


if __name__ == "__main__":
    main()
