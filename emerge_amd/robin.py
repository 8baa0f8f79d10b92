"""Robin / port surface assembly on boundary triangles, on the GPU.

Mirrors ``src/emerge/_emerge/physics/microwave/assembly/robinbc.py``:

* ``assemble_robin_bc(field, Bmat, surf_triangle_indices, gamma)`` (:434-445) -- adds
  the 8x8 blocks ``gamma * int f_i . f_j dA`` of the listed faces into the face-indexed
  array ``Bmat`` (``field.empty_tri_matrix()``, n_tris * 64);
* ``assemble_robin_bc_bvec(field, surf_triangle_indices, Ufunc)`` (:405-428) and
  ``assemble_robin_bc_bvec_scat`` (:576-611) -- the port excitation vector from the
  caller's mode function evaluated at the 6-point rule's points;
* ``RobinSurface`` -- the device-resident form: unit blocks computed once, then per
  frequency ``K += gamma(k0) * blocks`` added straight into the CSC values of the volume
  system (``K += field.generate_csc(B_matrix_robin)``, assembler.py:635), for the full
  matrix or the PEC-restricted block.

The kernels are in ``csrc/robin.hip`` (``em_ned2_robin_blocks``, ``em_ned2_robin_force``,
``em_tri_rowcol``, ``em_csc_map_find``, ``em_csc_add``). The mode function itself is the
caller's Python callable, evaluated on the host like in the reference.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check
from .assembly import CscMap, Ned2Plan, PREC_FP64, PREC_MIXED, _stream_ptr

# 6-point degree-4 triangle rule: weights, then barycentrics L1, L2, L3 of the points
# (robinbc.py:185-192).
DPTS = np.array([
    [0.10995174365532200, 0.10995174365532200, 0.10995174365532200, 0.22338158967801100,
     0.22338158967801100, 0.22338158967801100],
    [0.81684757298045896, 0.09157621350977101, 0.09157621350977101, 0.10810301816807000,
     0.44594849091596500, 0.44594849091596500],
    [0.09157621350977101, 0.81684757298045896, 0.09157621350977101, 0.44594849091596500,
     0.10810301816807000, 0.44594849091596500],
    [0.09157621350977004, 0.09157621350977008, 0.81684757298045807, 0.44594849091596500,
     0.44594849091596500, 0.10810301816807000],
], dtype=np.float64)


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def generate_points_3d(vertices: np.ndarray, tris: np.ndarray, surf: np.ndarray):
    """The rule's points on the listed faces, flattened point-major (index w * NS + i),
    as robinbc.py:111-135 returns them for the mode function."""
    surf = np.asarray(surf, dtype=np.int64)
    ids = tris[:, surf]
    out = []
    for q in range(3):
        c = vertices[q][ids]  # (3, NS)
        out.append((c[0][None, :] * DPTS[1][:, None] + c[1][None, :] * DPTS[2][:, None]
                    + c[2][None, :] * DPTS[3][:, None]).reshape(-1))
    return out[0], out[1], out[2]


class RobinSurface:
    """The listed boundary faces of one field on the device (a port or an absorbing
    boundary): their tables, unit blocks, force vectors and CSC positions."""

    def __init__(self, plan: Ned2Plan, field, surf_tris, stream=None):
        dev = plan.device
        self.plan = plan
        self.field = field
        self.n_tris = int(field.mesh.n_tris)
        self.surf_host = np.ascontiguousarray(np.asarray(surf_tris, dtype=np.int64).reshape(-1))
        self.n_surf = int(self.surf_host.shape[0])
        self.tris = torch.as_tensor(np.ascontiguousarray(field.mesh.tris, dtype=np.int64),
                                    device=dev)
        self.tri_to_field = torch.as_tensor(
            np.ascontiguousarray(field.tri_to_field, dtype=np.int64), device=dev)
        self.surf = torch.as_tensor(self.surf_host, device=dev)
        self.stream = stream

    def blocks(self, gamma: complex = 1.0, out: torch.Tensor | None = None,
               by_face: bool = False) -> torch.Tensor:
        """gamma * the 8x8 blocks, added into ``out`` (complex128; n_tris * 64 by face id
        when ``by_face``, else n_surf * 64 in list order; zeros when None)."""
        n = (self.n_tris if by_face else self.n_surf) * 64
        if out is None:
            out = torch.zeros(n, dtype=torch.complex128, device=self.plan.device)
        if out.dtype != torch.complex128 or out.numel() != n or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous complex128 tensor of {n} entries")
        g = complex(gamma)
        layout = _lib.EM_TRI_BY_FACE if by_face else _lib.EM_TRI_COMPACT
        check(_lib.lib().em_ned2_robin_blocks(self.plan._h, _vp(self.tris), self.n_tris,
                                              _vp(self.surf), self.n_surf, g.real, g.imag,
                                              _vp(out), layout, _stream_ptr(self.stream)))
        return out

    def force(self, U, U_curl=None, normals=None, bvec: torch.Tensor | None = None):
        """Add the port excitation of the mode field ``U`` (complex [3, 6, n_surf] at the
        rule's points) into ``bvec`` (complex128 [n_field]; zeros when None)."""
        dev = self.plan.device

        def cdev(a, shape):
            t = torch.as_tensor(np.asarray(a) if not isinstance(a, torch.Tensor) else a)
            t = t.to(device=dev, dtype=torch.complex128).reshape(shape).contiguous()
            return t
        shape = (3, 6, self.n_surf)
        Ut = cdev(U, shape)
        Ct = cdev(U_curl, shape) if U_curl is not None else None
        Nt = None
        if U_curl is not None:
            if normals is None:
                raise ValueError("the scattered-field form needs the normals")
            Nt = torch.as_tensor(np.asarray(normals, dtype=np.float64).reshape(3, self.n_surf),
                                 device=dev).contiguous()
        if bvec is None:
            bvec = torch.zeros(self.field.n_field, dtype=torch.complex128, device=dev)
        if bvec.dtype != torch.complex128 or bvec.numel() != self.field.n_field:
            raise ValueError("bvec must be complex128 with n_field entries")
        check(_lib.lib().em_ned2_robin_force(self.plan._h, _vp(self.tris), self.n_tris,
                                             _vp(self.tri_to_field), _vp(self.surf), self.n_surf,
                                             _vp(Ut), _vp(Ct), _vp(Nt), _vp(bvec),
                                             _stream_ptr(self.stream)))
        return bvec

    def rowcol(self):
        """rows/cols (int64, n_surf * 64) of the blocks, femdata.py:75-94 order."""
        n = self.n_surf * 64
        rows = torch.empty(max(n, 1), dtype=torch.int64, device=self.plan.device)
        cols = torch.empty(max(n, 1), dtype=torch.int64, device=self.plan.device)
        check(_lib.lib().em_tri_rowcol(_vp(self.tri_to_field), self.n_tris, _vp(self.surf),
                                       self.n_surf, _vp(rows), _vp(cols),
                                       _stream_ptr(self.stream)))
        return rows[:n], cols[:n]

    def positions(self, cscmap: CscMap, sel=None) -> torch.Tensor:
        """CSC positions of the block entries in ``cscmap`` (-1 where absent; with ``sel``
        the map is the restricted one and removed DOFs give -1)."""
        rows, cols = self.rowcol()
        pos = torch.empty_like(rows)
        ni = None if sel is None else sel.new_index
        check(_lib.lib().em_csc_map_find(cscmap._h, _vp(rows), _vp(cols), rows.numel(),
                                         _vp(ni), _vp(pos), _stream_ptr(self.stream)))
        return pos

    def add_to_csc(self, data_csc: torch.Tensor, pos: torch.Tensor, unit_blocks: torch.Tensor,
                   gamma: complex) -> torch.Tensor:
        """data_csc[pos] += gamma * unit_blocks (K += generate_csc(gamma * B))."""
        if data_csc.dtype not in (torch.complex128, torch.complex64):
            raise ValueError("CSC values must be complex128 or complex64")
        prec = PREC_FP64 if data_csc.dtype == torch.complex128 else PREC_MIXED
        g = complex(gamma)
        if unit_blocks.numel() != pos.numel() or unit_blocks.dtype != torch.complex128:
            raise ValueError("unit_blocks must be complex128 and match pos")
        check(_lib.lib().em_csc_add(_vp(data_csc), prec, _vp(pos), _vp(unit_blocks), pos.numel(),
                                    g.real, g.imag, _stream_ptr(self.stream)))
        return data_csc


def _plan(field, plan):
    if plan is not None:
        return plan
    from .curlcurl import plan_for
    return plan_for(field)


def assemble_robin_bc(field, Bmat: np.ndarray, surf_triangle_indices, gamma, plan=None):
    """robinbc.py:434-445: Bmat (n_tris * 64, complex128) with the blocks added."""
    rs = RobinSurface(_plan(field, plan), field, surf_triangle_indices)
    out = torch.as_tensor(np.ascontiguousarray(Bmat, dtype=np.complex128),
                          device=rs.plan.device).clone()
    rs.blocks(gamma, out=out, by_face=True)
    return out.cpu().numpy()


def assemble_robin_bc_bvec(field, surf_triangle_indices, Ufunc, plan=None):
    """robinbc.py:405-428: the excitation vector (complex128 [n_field])."""
    surf = np.asarray(surf_triangle_indices, dtype=np.int64)
    x, y, z = generate_points_3d(field.mesh.nodes, field.mesh.tris, surf)
    U = np.asarray(Ufunc(x, y, z)).reshape(3, DPTS.shape[1], surf.shape[0])
    rs = RobinSurface(_plan(field, plan), field, surf)
    return rs.force(U).cpu().numpy()


def assemble_robin_bc_bvec_scat(field, surf_triangle_indices, Ufunc, UfuncCurl, normals,
                                plan=None):
    """robinbc.py:576-611: the scattered-field excitation vector."""
    surf = np.asarray(surf_triangle_indices, dtype=np.int64)
    x, y, z = generate_points_3d(field.mesh.nodes, field.mesh.tris, surf)
    U = np.asarray(Ufunc(x, y, z)).reshape(3, DPTS.shape[1], surf.shape[0])
    C = np.asarray(UfuncCurl(x, y, z)).reshape(3, DPTS.shape[1], surf.shape[0])
    rs = RobinSurface(_plan(field, plan), field, surf)
    return rs.force(U, C, normals).cpu().numpy()
