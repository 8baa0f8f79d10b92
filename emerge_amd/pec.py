"""PEC elimination and the hand-off of the assembled system to a sparse solver, on the GPU.

Mirrors the steps that follow the assembly in the reference:

* ``pec_solve_ids`` -- the prescribed (PEC) DOFs of ``Assembler.assemble_freq_matrix``
  (``physics/microwave/assembly/assembler.py:533-564``: all 20 DOFs of every conductor tet;
  for every PEC boundary face the DOFs of its edges, ``edge_to_field``, and its
  ``tri_to_field`` column) and ``solve_ids``, the ascending complement (``:742``);
* ``select_csc`` / ``select_system`` -- ``Asel = A[:, solve_ids][solve_ids, :]``,
  ``bsel = b[solve_ids, :]`` (``solver.py:2071-2072``; the eigen path's
  ``A[np.ix_(solve_ids, solve_ids)]``, ``solver.py:2229-2230``, is the same block), with
  ``K.eliminate_zeros()`` (``assembler.py:760``) as the ``drop_zeros`` option;
* ``scatter_solution`` -- ``solution[solve_ids, :] = x`` (``solver.py:2118-2120``);
* ``CscMap.restrict`` (in ``assembly``) -- the fused form for sweeps: the free-DOF block
  of ``to_csc(Evec - Bvec*K0**2)`` straight from the COO values.

Every step runs in libemerge_amd (``em_ned2_pec_mask``, ``em_dof_compact``,
``em_csc_select_*``, ``em_dof_gather``/``em_dof_scatter``); there is no host fallback.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import check
from .assembly import PREC_FP64, PREC_MIXED, Ned2Plan, _stream_ptr, to_host


def _vp(t: torch.Tensor | None) -> ctypes.c_void_p | None:
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _dev_i64(a, device) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.int64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=np.int64)).reshape(-1),
                           device=device)


@dataclass
class SolveIds:
    """The free DOFs of a system: ``keep`` (uint8 [N]), ``solve_ids`` (int64 [n_solve],
    ascending) and ``new_index`` (int64 [N]: position in ``solve_ids`` or -1), on the
    device."""
    keep: torch.Tensor
    solve_ids: torch.Tensor
    new_index: torch.Tensor
    n_field: int

    @property
    def n_solve(self) -> int:
        return int(self.solve_ids.numel())

    @classmethod
    def from_keep(cls, keep: torch.Tensor, stream=None) -> "SolveIds":
        """Compact a uint8/bool keep mask of N DOFs (``em_dof_compact``)."""
        keep = keep.to(torch.uint8).contiguous()
        if not keep.is_cuda:
            raise ValueError("keep mask must be on the GPU")
        n = keep.numel()
        new_index = torch.empty(n, dtype=torch.int64, device=keep.device)
        ids = torch.empty(max(n, 1), dtype=torch.int64, device=keep.device)
        ns = ctypes.c_int64()
        check(_lib.lib().em_dof_compact(n, _vp(keep), _vp(new_index), _vp(ids), ctypes.byref(ns),
                                        _stream_ptr(stream)))
        return cls(keep, ids[:ns.value], new_index, n)


def pec_solve_ids(plan: Ned2Plan, field=None, conductor_tets=(), pec_tris=(),
                  tri_to_field=None, stream=None) -> SolveIds:
    """Free DOFs after removing the PEC DOFs (assembler.py:533-564, 742).

    ``conductor_tets`` are plan tet ids; ``pec_tris`` face ids of ``field`` (the faces of
    the PEC boundary tags, ``mesh.get_triangles(face_tags)`` in the reference). The face
    DOF table is ``field.tri_to_field`` unless ``tri_to_field`` (int64 [8, F]) is given.
    """
    dev = plan.device
    cond = _dev_i64(conductor_tets, dev)
    tris = _dev_i64(pec_tris, dev)
    if tri_to_field is None:
        if tris.numel() and field is None:
            raise ValueError("PEC faces need the field (or its tri_to_field table)")
        tri_to_field = field.tri_to_field if field is not None else np.zeros((8, 0), np.int64)
    t2f = _dev_i64(tri_to_field, dev)
    tt = np.asarray(tri_to_field) if not isinstance(tri_to_field, torch.Tensor) else tri_to_field
    if tuple(tt.shape)[0] != 8:
        raise ValueError("tri_to_field must have 8 rows")
    n_tris = int(tt.shape[1])
    keep = torch.empty(plan.n_field, dtype=torch.uint8, device=dev)
    check(_lib.lib().em_ned2_pec_mask(plan._h, _vp(cond), cond.numel(), _vp(t2f), n_tris,
                                      _vp(tris), tris.numel(), _vp(keep), _stream_ptr(stream)))
    return SolveIds.from_keep(keep, stream)


def select_csc(indptr: torch.Tensor, indices: torch.Tensor, data: torch.Tensor | None,
               sel: SolveIds, drop_zeros: bool = False, stream=None):
    """``A[:, solve_ids][solve_ids, :]`` of a device CSC matrix with sorted rows.

    Returns device ``(indptr, indices, data)`` of the ``n_solve x n_solve`` block
    (``data`` None when ``data`` is None). ``drop_zeros`` also removes explicit zeros.
    """
    n = indptr.numel() - 1
    if n != sel.n_field:
        raise ValueError(f"matrix has {n} columns, solve ids are for {sel.n_field}")
    if drop_zeros and data is None:
        raise ValueError("drop_zeros needs the values")
    prec = PREC_MIXED if (data is not None and data.dtype == torch.complex64) else PREC_FP64
    if data is not None and data.dtype not in (torch.complex128, torch.complex64):
        raise ValueError("values must be complex128 or complex64")
    indptr, indices = indptr.contiguous(), indices.contiguous()
    data = data.contiguous() if data is not None else None
    L = _lib.lib()
    sp = torch.empty(sel.n_solve + 1, dtype=torch.int64, device=indptr.device)
    nz = ctypes.c_int64()
    check(L.em_csc_select_count(n, _vp(indptr), _vp(indices), _vp(data), prec, int(drop_zeros),
                                _vp(sel.new_index), sel.n_solve, _vp(sp), ctypes.byref(nz),
                                _stream_ptr(stream)))
    si = torch.empty(max(nz.value, 1), dtype=torch.int64, device=indptr.device)
    sd = None
    if data is not None:
        sd = torch.empty(max(nz.value, 1), dtype=data.dtype, device=indptr.device)
    check(L.em_csc_select_fill(n, _vp(indptr), _vp(indices), _vp(data), prec, int(drop_zeros),
                               _vp(sel.new_index), _vp(sp), _vp(si), _vp(sd),
                               _stream_ptr(stream)))
    return sp, si[:nz.value], (sd[:nz.value] if sd is not None else None)


def _move(ids: torch.Tensor, src: torch.Tensor, dst: torch.Tensor, scatter: bool, stream):
    prec = PREC_MIXED if src.dtype == torch.complex64 else PREC_FP64
    if src.dtype != dst.dtype or src.dtype not in (torch.complex128, torch.complex64):
        raise ValueError("vectors must share a complex dtype")
    ncols = 1 if src.dim() == 1 else src.shape[1]
    fn = _lib.lib().em_dof_scatter if scatter else _lib.lib().em_dof_gather
    check(fn(_vp(ids), ids.numel(), ncols, _vp(src), _vp(dst), prec, _stream_ptr(stream)))
    return dst


def gather_rows(b: torch.Tensor, sel: SolveIds, stream=None) -> torch.Tensor:
    """``b[solve_ids, :]`` (solver.py:2072) of a device vector or row-major matrix."""
    b = b.contiguous()
    if b.shape[0] != sel.n_field:
        raise ValueError(f"b has {b.shape[0]} rows, expected {sel.n_field}")
    out = torch.empty((sel.n_solve,) + tuple(b.shape[1:]), dtype=b.dtype, device=b.device)
    return _move(sel.solve_ids, b, out, False, stream)


def scatter_solution(x: torch.Tensor, sel: SolveIds, stream=None) -> torch.Tensor:
    """``solution = zeros((N, NB)); solution[solve_ids, :] = x`` (solver.py:2118-2120)."""
    x = x.contiguous()
    if x.shape[0] != sel.n_solve:
        raise ValueError(f"x has {x.shape[0]} rows, expected {sel.n_solve}")
    out = torch.zeros((sel.n_field,) + tuple(x.shape[1:]), dtype=x.dtype, device=x.device)
    return _move(sel.solve_ids, x, out, True, stream)


def select_system(A, b, sel: SolveIds, drop_zeros: bool = False, stream=None):
    """``Asel, bsel`` of ``SolveRoutine.solve`` (solver.py:2071-2072).

    ``A`` is a device triple ``(indptr, indices, data)`` or a SciPy CSC matrix (copied to
    the device); ``b`` a vector or ``(N, NB)`` matrix (host or device). Returns
    ``(indptr, indices, data)`` and ``bsel`` on the device.
    """
    dev = sel.keep.device
    if not isinstance(A, tuple):
        A = A.tocsc()
        if not A.has_sorted_indices:
            A = A.sorted_indices()
        A = (torch.as_tensor(A.indptr.astype(np.int64), device=dev),
             torch.as_tensor(A.indices.astype(np.int64), device=dev),
             torch.as_tensor(np.asarray(A.data, dtype=np.complex128), device=dev))
    Asel = select_csc(*A, sel, drop_zeros=drop_zeros, stream=stream)
    bt = b if isinstance(b, torch.Tensor) else torch.as_tensor(np.asarray(b, dtype=np.complex128))
    bsel = gather_rows(bt.to(dev), sel, stream)
    return Asel, bsel


def to_scipy(indptr: torch.Tensor, indices: torch.Tensor, data: torch.Tensor, n: int):
    """SciPy ``csc_matrix`` of a device CSC triple (host copy)."""
    import scipy.sparse as sp
    return sp.csc_matrix((to_host(data), to_host(indices), to_host(indptr)), shape=(n, n))
