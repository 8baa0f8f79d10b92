"""S-parameters of modal ports from a solved field.

Mirrors src/emerge/_emerge/physics/microwave/sparam.py:26-97 and the modal branch of
``Microwave3D._compute_s_data`` (physics/microwave/microwave_3d.py:2553-2571): the
field-mode overlap ``int (E - Q E_mode) . conj(E_mode) / (2 Z0) dA`` over the port faces
(6-point rule, order 5) and the mode power ``int |E_mode|^2 / (2 Z0) dA`` (7-point rule,
order 7), with ``S_ij = overlap / mode power * |port power|`` (microwave_3d.py:2364-2368).
The surface integral follows mth/integrals.py:31-42 / :64-104. The field at the rule's
points comes from the GPU interpolation (``emerge_amd.interp``); the mode field is the
port's own (host) function, as in the reference.
"""
from __future__ import annotations

import numpy as np

from .ports import Z0
from .robin import DPTS


def _radon7() -> np.ndarray:
    """Radon's 7-point degree-5 triangle rule (the reference's order-7 table,
    mth/strang_rules.py STRANG7): weights, then L1, L2, L3."""
    r = np.sqrt(15.0)
    a1, a2 = (6 - r) / 21, (6 + r) / 21
    w1, w2 = (155 - r) / 1200, (155 + r) / 1200
    pts = [(1 / 3, 1 / 3, 1 / 3)]
    pts += [(1 - 2 * a1, a1, a1), (a1, 1 - 2 * a1, a1), (a1, a1, 1 - 2 * a1)]
    pts += [(1 - 2 * a2, a2, a2), (a2, 1 - 2 * a2, a2), (a2, a2, 1 - 2 * a2)]
    W = [0.225] + [w1] * 3 + [w2] * 3
    L = np.array(pts).T
    return np.vstack([np.array(W)[None, :], L])


def gaus_quad_tri(p: int) -> np.ndarray:
    """mth/optimized.py:30-55 for the two orders the S-parameters use."""
    if p == 5:
        return DPTS.copy()
    if p == 7:
        return _radon7()
    raise ValueError(f"triangle rule of order {p} is not provided")


def generate_int_points_tri(nodes, triangles, PTS):
    """mth/optimized.py:88-118: points (rule-point-major) and the (npts, ntri) shape."""
    c = nodes[:, triangles]  # (3 coords, 3 vertices, ntri)
    pts = [c[q, 0][None, :] * PTS[1][:, None] + c[q, 1][None, :] * PTS[2][:, None]
           + c[q, 2][None, :] * PTS[3][:, None] for q in range(3)]
    return pts[0].ravel(), pts[1].ravel(), pts[2].ravel(), (PTS.shape[1], triangles.shape[1])


def surface_integral(nodes, triangles, function, constants=None, gq_order: int = 4):
    """mth/integrals.py:64-104 with :31-42: sum over faces of C * A * sum_w w f(x_w)."""
    triangles = np.asarray(triangles, np.int64)
    if constants is None:
        constants = np.ones(triangles.shape[1])
    D = gaus_quad_tri(gq_order)
    x, y, z, shape = generate_int_points_tri(nodes, triangles, D)
    fA = np.asarray(function(x, y, z)).reshape(shape)
    v = nodes[:, triangles]
    A = 0.5 * np.linalg.norm(np.cross((v[:, 1] - v[:, 0]).T, (v[:, 2] - v[:, 0]).T), axis=1)
    return np.sum(np.asarray(constants) * np.sum(D[0][:, None] * fA, axis=0) * A)


def sparam_mode_power(nodes, tri_vertices, bc, k0, const, gq_order: int = 7):
    """sparam.py:26-56."""
    def inproduct2(x, y, z):
        E = bc.port_mode_3d_global(x, y, z, k0, which="Exy")
        return np.sum(E * np.conj(E), axis=0) / (2 * Z0)
    return surface_integral(nodes, tri_vertices, inproduct2, const, gq_order=gq_order)


def sparam_field_power(nodes, tri_vertices, bc, active: bool, k0, fieldf, const,
                       gq_order: int = 5):
    """sparam.py:59-97."""
    Q = 1 if active else 0

    def inproduct1(x, y, z):
        mode = bc.port_mode_3d_global(x, y, z, k0, which="Exy")
        E1 = np.asarray(fieldf(x, y, z)) - Q * mode
        return np.sum(E1 * np.conj(mode), axis=0) / (2 * Z0)
    return surface_integral(nodes, tri_vertices, inproduct1, const, gq_order=gq_order)


def s_parameter(nodes, port_tris_vertices, bc, active: bool, k0, fieldf,
                port_power: float = 1.0):
    """S_ij of the passive port ``bc`` for the field driven at the active port
    (microwave_3d.py:2349-2368 with the modal branch of _compute_s_data)."""
    const = np.ones(port_tris_vertices.shape[1], np.complex128)
    field_p = sparam_field_power(nodes, port_tris_vertices, bc, active, k0, fieldf, const, 5)
    mode_p = sparam_mode_power(nodes, port_tris_vertices, bc, k0, const, 7)
    return field_p / mode_p * abs(port_power)
