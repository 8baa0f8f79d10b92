"""Host side of the Nedelec-2 assembly: device plans and material descriptors.

``Ned2Plan`` keeps one mesh resident on the GPU (through the C-ABI in
include/emerge_amd.h) and launches the HIP kernels; torch tensors are only the device
buffers. The reference-shaped entry point ``tet_mass_stiffness_matrices`` lives in
``emerge_amd.curlcurl``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import check, em_material_t

PREC_FP64 = _lib.EM_PREC_FP64
PREC_MIXED = _lib.EM_PREC_MIXED
PREC_FP64_REF = _lib.EM_PREC_FP64_REF
_PREC = {"fp64": PREC_FP64, "mixed": PREC_MIXED, "fp64_ref": PREC_FP64_REF,
         PREC_FP64: PREC_FP64, PREC_MIXED: PREC_MIXED, PREC_FP64_REF: PREC_FP64_REF}


def to_host(t: torch.Tensor) -> np.ndarray:
    """A fresh NumPy copy of a device tensor, through pinned host memory: on one MI355X a
    3.5 GB complex128 COO array takes ~0.88 s into pageable memory (4.0 GB/s, page faults
    of the fresh array included) and ~0.09 s into pinned memory (39 GB/s). The array
    keeps its pinned tensor alive; torch's host allocator reuses the block once both are
    freed."""
    h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    h.copy_(t, non_blocking=True)
    torch.cuda.current_stream(t.device).synchronize()
    return h.numpy()


def _prec(p) -> int:
    """Precision name or code: ``"fp64"`` (exact element integrals, complex128),
    ``"mixed"`` (FP32 contraction, complex64) or ``"fp64_ref"`` (the reference's own
    11-point arithmetic in absolute coordinates, complex128; assembly only)."""
    try:
        return _PREC[p]
    except KeyError:
        raise ValueError(f"unknown precision {p!r}; use 'fp64', 'mixed' or 'fp64_ref'") from None


def _out_dtype(prec: int):
    return torch.complex64 if prec == PREC_MIXED else torch.complex128


def _stream_ptr(stream) -> ctypes.c_void_p:
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a), dtype=np.int64)


@dataclass
class Material:
    """A complex128 device tensor field and how to index it (see ``em_material_t``).

    ``tensor`` keeps the device storage alive for as long as the descriptor is used.
    """
    tensor: torch.Tensor
    kind: int
    tet_stride: int
    comp_stride: int
    flags: int = 0

    def c_struct(self) -> em_material_t:
        return em_material_t(ctypes.c_void_p(self.tensor.data_ptr()), self.kind, self.flags,
                             self.tet_stride, self.comp_stride)


_OFF_DIAG = ([0, 0, 1, 1, 2, 2], [1, 2, 0, 2, 0, 1])


def ref_structure(a) -> tuple[bool, bool]:
    """For a (3, 3, T) array (NumPy, or a torch tensor on any device, checked where it
    lives): (every off-diagonal entry is zero, the three diagonal entries are equal)."""
    if isinstance(a, torch.Tensor):
        # per-component views: no copy of the tensor, one synchronisation per question
        nz = torch.stack([(a[p, q] != 0).any() for p, q in zip(*_OFF_DIAG)])
        off_zero = not bool(nz.any())
        same = torch.stack([(a[0, 0] != a[1, 1]).any(), (a[0, 0] != a[2, 2]).any()])
        return off_zero, not bool(same.any())
    off_zero = not any(np.any(a[p, q]) for p, q in zip(*_OFF_DIAG))
    return off_zero, bool(np.array_equal(a[0, 0], a[1, 1]) and np.array_equal(a[0, 0], a[2, 2]))


def material(value, n_tets: int, device, inverse: bool = False, detect: bool = True,
             layout: str | None = None) -> Material:
    """Describe a per-tet material tensor field for the kernels.

    Layouts: ``"ref"`` is the reference's ``(3, 3, T)`` array (``assembler.py:440-453``);
    ``"tet"`` is tet-major: ``(T,)`` scalar per tet, ``(T, 3)`` diagonal, ``(T, 3, 3)``
    full. A Python/NumPy scalar is one value for every tet. By default NumPy arrays of
    shape ``(3, 3, T)`` are ``"ref"`` and everything else is ``"tet"``.
    For a ``"ref"`` array with ``detect=True`` the structure is checked (on the host for a
    NumPy array, on the device for a tensor there: one reduction and a synchronisation)
    and a diagonal or scalar field is described as such (its diagonal is read in place).
    ``inverse`` marks a mu field that already holds mu_r^-1.
    """
    flags = _lib.EM_MAT_IS_INVERSE if inverse else 0
    if np.isscalar(value) or (isinstance(value, (np.ndarray, torch.Tensor)) and value.ndim == 0):
        t = torch.tensor([complex(value)], dtype=torch.complex128, device=device)
        return Material(t, _lib.EM_MAT_SCALAR, 0, 0, flags)
    is_torch = isinstance(value, torch.Tensor)
    shape = tuple(value.shape)
    if layout is None:
        layout = "ref" if (not is_torch and len(shape) == 3 and shape[:2] == (3, 3)) else "tet"
    if is_torch:
        t = value.to(device=device, dtype=torch.complex128).contiguous()
    else:
        arr = np.ascontiguousarray(value, dtype=np.complex128)
        t = torch.from_numpy(arr).to(device)
    if layout == "ref":
        if shape != (3, 3, n_tets):
            raise ValueError(f"reference-layout material must be (3, 3, {n_tets}), got {shape}")
        kind = _lib.EM_MAT_FULL
        if detect:
            off_zero, diag_equal = ref_structure(t if is_torch else arr)
            if off_zero:
                kind = _lib.EM_MAT_SCALAR if diag_equal else _lib.EM_MAT_DIAG
        if kind == _lib.EM_MAT_FULL:
            return Material(t, kind, 1, n_tets, flags)
        if kind == _lib.EM_MAT_DIAG:
            return Material(t, kind, 1, 4 * n_tets, flags)
        return Material(t, kind, 1, 0, flags)
    if shape == (n_tets,):
        return Material(t, _lib.EM_MAT_SCALAR, 1, 0, flags)
    if shape == (n_tets, 3):
        return Material(t, _lib.EM_MAT_DIAG, 3, 1, flags)
    if shape == (n_tets, 3, 3):
        return Material(t, _lib.EM_MAT_FULL, 9, 1, flags)
    raise ValueError(f"material of shape {shape} does not match {n_tets} tets")


class Ned2Plan:
    """A mesh resident on one GPU, ready for repeated assembly.

    Parameters follow the reference tables (``Mesh3D``/``Nedelec2`` attributes):
    ``nodes (3,Nn)``, ``tets (4,T)``, ``tris (3,F)``, ``edges (2,E)``,
    ``tet_to_field (20,T)``, ``tet_to_edge (6,T)``, ``tet_to_tri (4,T)``, ``n_field``,
    optional ``tetids`` (the assembled tets, conductor tets removed) and ``dofcodes``.
    """

    def __init__(self, nodes, tets, tris, edges, tet_to_field, tet_to_edge, tet_to_tri,
                 n_field: int, tetids=None, dofcodes=None, device=None):
        L = _lib.lib()
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("Ned2Plan needs a GPU device")
        nodes = np.ascontiguousarray(np.asarray(nodes), dtype=np.float64)
        tets = _i64(tets)
        tris, edges = _i64(tris), _i64(edges)
        t2f, t2e, t2t = _i64(tet_to_field), _i64(tet_to_edge), _i64(tet_to_tri)
        if dofcodes is None:
            from .mesh import DOFCODES_NED2
            dofcodes = DOFCODES_NED2
        dofcodes = _i64(dofcodes)
        self.n_tets = int(tets.shape[1])
        if tetids is not None:
            tetids = _i64(tetids)
            n_assy = int(tetids.shape[0])
        else:
            n_assy = self.n_tets
        self._keep = (nodes, tets, tris, edges, t2f, t2e, t2t, tetids, dofcodes)
        h = ctypes.c_void_p()
        ptr = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        check(L.em_ned2_plan_create(ctypes.byref(h), self.device.index or 0, ptr(nodes),
                                    nodes.shape[1], ptr(tets), self.n_tets, ptr(tris),
                                    tris.shape[1], ptr(edges), edges.shape[1], ptr(t2f), ptr(t2e),
                                    ptr(t2t), int(n_field), ptr(tetids), n_assy, ptr(dofcodes),
                                    int(dofcodes.shape[0])))
        self._h = h
        self._keep = None
        self.n_assy = n_assy
        self.n_field = int(n_field)
        self.nnz_coo = 400 * n_assy
        self.tet_offsets = np.array([0, self.n_tets], np.int64)
        self.dof_offsets = np.array([0, self.n_field], np.int64)

    @classmethod
    def from_field(cls, field, tetids=None, device=None) -> "Ned2Plan":
        m = field.mesh
        return cls(m.nodes, m.tets, m.tris, m.edges, field.get_tet_to_field(), m.tet_to_edge,
                   m.tet_to_tri, field.n_field, tetids=tetids, dofcodes=field.dofcodes3d,
                   device=device)

    @classmethod
    def from_fields(cls, fields, device=None) -> "Ned2Plan":
        """One plan over several meshes (a batch of uneven size), assembled in one launch.

        The meshes are concatenated with their node, edge, face and DOF numbers offset,
        which keeps every mesh's vertex order (so every element block is unchanged) and
        makes the global system block diagonal. ``plan.tet_offsets[i]`` is the first
        assembly index of mesh i (its COO values start at 400 * that), and
        ``plan.dof_offsets[i]`` the first global DOF of mesh i.
        """
        nodes, tets, tris, edges, t2f, t2e, t2t = [], [], [], [], [], [], []
        nn = ne = nf_ = ndof = 0
        toffs, doffs = [0], [0]
        for f in fields:
            m = f.mesh
            nodes.append(m.nodes)
            tets.append(m.tets + nn)
            tris.append(m.tris + nn)
            edges.append(m.edges + nn)
            t2e.append(m.tet_to_edge + ne)
            t2t.append(m.tet_to_tri + nf_)
            t2f.append(f.get_tet_to_field() + ndof)
            nn += m.n_nodes
            ne += m.n_edges
            nf_ += m.n_tris
            ndof += f.n_field
            toffs.append(toffs[-1] + m.n_tets)
            doffs.append(ndof)
        plan = cls(np.concatenate(nodes, axis=1), np.concatenate(tets, axis=1),
                   np.concatenate(tris, axis=1), np.concatenate(edges, axis=1),
                   np.concatenate(t2f, axis=1), np.concatenate(t2e, axis=1),
                   np.concatenate(t2t, axis=1), ndof, device=device)
        plan.tet_offsets = np.asarray(toffs, np.int64)
        plan.dof_offsets = np.asarray(doffs, np.int64)
        return plan

    def set_tuning(self, tile: int) -> None:
        """Tets per workgroup tile of the assembly kernels (32, 64 or 128)."""
        check(_lib.lib().em_ned2_plan_set_tuning(self._h, int(tile)))

    def get_tuning(self) -> int:
        v = ctypes.c_int32()
        check(_lib.lib().em_ned2_plan_get_tuning(self._h, ctypes.byref(v)))
        return v.value

    def set_window(self, window: int) -> None:
        """Tets per window of the tile order (rebuilds the tile lists; synchronises)."""
        check(_lib.lib().em_ned2_plan_set_window(self._h, int(window)))

    def get_window(self) -> int:
        v = ctypes.c_int64()
        check(_lib.lib().em_ned2_plan_get_window(self._h, ctypes.byref(v)))
        return v.value

    def set_prefetch(self, chunk_tets: int) -> None:
        """Prefetched launches: chunks of about ``chunk_tets`` tets, each launched right
        after a kernel that reads its inputs into the memory-side cache (0: one launch).
        Same outputs bit for bit; set by the autotuner. A new plan starts with 2**20 if it
        assembles at least 2**21 tets, else 0."""
        check(_lib.lib().em_ned2_plan_set_prefetch(self._h, int(chunk_tets)))

    def get_prefetch(self) -> int:
        v = ctypes.c_int64()
        check(_lib.lib().em_ned2_plan_get_prefetch(self._h, ctypes.byref(v)))
        return v.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().em_ned2_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- materials ------------------------------------------------------------
    def material(self, value, inverse=False, detect=True, layout=None) -> Material:
        if isinstance(value, Material):
            return value
        return material(value, self.n_tets, self.device, inverse=inverse, detect=detect,
                        layout=layout)

    # -- launches -------------------------------------------------------------
    def assemble_km(self, er, ur, precision="fp64", out=None, stream=None):
        """K (stiffness, mu_r^-1) and M (mass, eps_r) COO value arrays, reference layout."""
        prec = _prec(precision)
        er_m, ur_m = self.material(er), self.material(ur)
        dt = _out_dtype(prec)
        if out is None:
            K = torch.empty(self.nnz_coo, dtype=dt, device=self.device)
            M = torch.empty(self.nnz_coo, dtype=dt, device=self.device)
        else:
            K, M = out
            _check_out(K, self.nnz_coo, dt)
            _check_out(M, self.nnz_coo, dt)
        e, u = er_m.c_struct(), ur_m.c_struct()
        check(_lib.lib().em_ned2_assemble_km(self._h, ctypes.byref(e), ctypes.byref(u),
                                             ctypes.c_void_p(K.data_ptr()),
                                             ctypes.c_void_p(M.data_ptr()), prec,
                                             _stream_ptr(stream)))
        return K, M

    def assemble_a(self, er, ur, k0, precision="fp64", out=None, stream=None):
        """COO values of A = K - k0^2 M. ``k0`` a float (returns ``[nnz_coo]``) or a
        sequence of wavenumbers (returns ``[len(k0), nnz_coo]``, one launch)."""
        prec = _prec(precision)
        er_m, ur_m = self.material(er), self.material(ur)
        dt = _out_dtype(prec)
        scalar = np.isscalar(k0)
        k0a = np.ascontiguousarray(np.atleast_1d(np.asarray(k0, dtype=np.float64)))
        nf = int(k0a.shape[0])
        if out is None:
            out = torch.empty((nf, self.nnz_coo), dtype=dt, device=self.device)
        else:
            _check_out(out, nf * self.nnz_coo, dt, rows=nf)
        stride = self.nnz_coo if out.dim() == 1 else out.stride(0)
        e, u = er_m.c_struct(), ur_m.c_struct()
        check(_lib.lib().em_ned2_assemble_a(self._h, ctypes.byref(e), ctypes.byref(u),
                                            k0a.ctypes.data_as(ctypes.c_void_p), nf,
                                            ctypes.c_void_p(out.data_ptr()), stride, prec,
                                            _stream_ptr(stream)))
        if scalar and out.dim() == 2:
            return out[0]
        return out

    def assemble_csc(self, cscmap, er, ur, k0: float = 0.0, which: str = "A",
                     precision="fp64", out=None, stream=None) -> torch.Tensor:
        """CSC values (in ``cscmap``'s pattern) of ``A = K - k0^2 M`` (``which="A"``), of K
        (``"K"``) or of M (``"M"``), assembled without the COO arrays
        (``em_ned2_assemble_csc``); equal bit for bit to ``to_csc`` of ``assemble_a`` /
        ``assemble_km``. ``cscmap`` is a map of this plan or restricted from one."""
        prec = _prec(precision)
        if prec == _lib.EM_PREC_FP64_REF:
            raise ValueError("assemble_csc takes 'fp64' or 'mixed'")
        w = {"A": 0, "K": 1, "M": 2}[which]
        er_m, ur_m = self.material(er), self.material(ur)
        dt = _out_dtype(prec)
        if out is None:
            out = torch.empty(cscmap.nnz, dtype=dt, device=self.device)
        else:
            _check_out(out, cscmap.nnz, dt)
        e, u = er_m.c_struct(), ur_m.c_struct()
        check(_lib.lib().em_ned2_assemble_csc(self._h, cscmap._h, ctypes.byref(e),
                                              ctypes.byref(u), float(k0), w,
                                              ctypes.c_void_p(out.data_ptr()), prec,
                                              _stream_ptr(stream)))
        return out

    def backward(self, grad_coo: torch.Tensor, k0: float, eps_kind: int | None, grad_eps,
                 mu_kind: int | None, grad_mu_inv, stream=None) -> None:
        """dL/d eps_r and dL/d mu_r^-1 (tet-major complex128 buffers, written in place)
        from dL/dA of the COO values of ``assemble_a`` at wavenumber ``k0``."""
        if grad_coo.dtype != torch.complex128 or grad_coo.numel() != self.nnz_coo:
            raise ValueError(f"upstream gradient must be complex128 with {self.nnz_coo} entries")
        grad_coo = grad_coo.contiguous()
        width = {_lib.EM_MAT_SCALAR: 1, _lib.EM_MAT_DIAG: 3, _lib.EM_MAT_FULL: 9}
        ptrs = []
        for kind, g in ((eps_kind, grad_eps), (mu_kind, grad_mu_inv)):
            if g is None:
                ptrs.append(None)
                continue
            if g.dtype != torch.complex128 or not g.is_contiguous() or \
                    g.numel() != self.n_tets * width[kind]:
                raise ValueError("gradient buffers must be contiguous complex128 of the leaf size")
            ptrs.append(ctypes.c_void_p(g.data_ptr()))
        check(_lib.lib().em_ned2_backward(self._h, ctypes.c_void_p(grad_coo.data_ptr()), float(k0),
                                          int(eps_kind or 0), ptrs[0], int(mu_kind or 0), ptrs[1],
                                          _stream_ptr(stream)))

    def rowcol(self, stream=None):
        rows = torch.empty(self.nnz_coo, dtype=torch.int64, device=self.device)
        cols = torch.empty(self.nnz_coo, dtype=torch.int64, device=self.device)
        check(_lib.lib().em_ned2_rowcol(self._h, ctypes.c_void_p(rows.data_ptr()),
                                        ctypes.c_void_p(cols.data_ptr()), _stream_ptr(stream)))
        return rows, cols


class CscMap:
    """Device CSC pattern of a plan's COO matrix (``CSCMapping``, csc_cast.py:29-63).

    Built on the GPU from the DOF map (``em_csc_map_create``). ``to_csc`` sums the COO
    values of every CSC entry on the GPU.
    """

    def __init__(self, plan: Ned2Plan, stream=None, *, _handle=None, _device=None):
        L = _lib.lib()
        if _handle is None:
            h = ctypes.c_void_p()
            check(L.em_csc_map_create(ctypes.byref(h), plan._h, _stream_ptr(stream)))
            self.device = plan.device
        else:
            h = _handle
            self.device = _device
        self._h = h
        # a restricted map (``restrict``) covers the free-DOF block only
        self.restricted = _handle is not None
        n, nnz, nnz_coo = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(L.em_csc_map_sizes(h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(nnz_coo)))
        self.N, self.nnz, self.nnz_coo = n.value, nnz.value, nnz_coo.value
        self._cache = {}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().em_csc_map_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def export(self, *names, stream=None) -> dict:
        """Device int64 tensors among indptr, indices, perm, target_csc, col_offsets."""
        sizes = {"indptr": self.N + 1, "indices": self.nnz, "perm": self.nnz_coo,
                 "target_csc": self.nnz_coo, "col_offsets": self.N + 1}
        out = {}
        ptrs = []
        for k in ("indptr", "indices", "perm", "target_csc", "col_offsets"):
            if k in names:
                t = torch.empty(sizes[k], dtype=torch.int64, device=self.device)
                out[k] = t
                ptrs.append(ctypes.c_void_p(t.data_ptr()))
            else:
                ptrs.append(None)
        check(_lib.lib().em_csc_map_export(self._h, *ptrs, _stream_ptr(stream)))
        return out

    def restrict(self, sel, stream=None) -> "CscMap":
        """Map of the free-DOF block ``A[:, solve_ids][solve_ids, :]`` (solver.py:2071):
        its ``to_csc``/``to_csc_combine`` read the same COO values and write the block's
        values directly. ``sel`` is a ``SolveIds`` of this map's DOFs."""
        if self.restricted:
            raise ValueError("map is already restricted")
        if sel.n_field != self.N:
            raise ValueError(f"solve ids are for {sel.n_field} DOFs, map has {self.N}")
        h = ctypes.c_void_p()
        check(_lib.lib().em_csc_map_restrict(self._h, ctypes.c_void_p(sel.new_index.data_ptr()),
                                             sel.n_solve, ctypes.byref(h), _stream_ptr(stream)))
        return CscMap(None, _handle=h, _device=self.device)

    def to_csc(self, data: torch.Tensor, out=None, stream=None,
               block_symmetric: bool = False, segment_gather: bool = False) -> torch.Tensor:
        """CSC values of the COO values ``data`` (complex128 or complex64 on device).

        ``block_symmetric``: every 20x20 block of ``data`` is symmetric (every block the
        element kernels write is, whatever the tensors: the reference mirrors i <= j,
        curlcurl.py:357-371); selects the mirrored-read gather, same result, faster. ``segment_gather`` selects the per-entry perm/seg gather instead of
        the per-column incidence gather (same result)."""
        prec = PREC_FP64 if data.dtype == torch.complex128 else PREC_MIXED
        if data.dtype not in (torch.complex128, torch.complex64) or data.numel() != self.nnz_coo:
            raise ValueError(f"COO data must be complex with {self.nnz_coo} entries")
        data = data.contiguous()
        if out is None:
            out = torch.empty(self.nnz, dtype=data.dtype, device=self.device)
        _check_out(out, self.nnz, data.dtype)
        flags = ((_lib.EM_CSC_BLOCK_SYMMETRIC if block_symmetric else 0)
                 | (_lib.EM_CSC_SEGMENT_GATHER if segment_gather else 0))
        check(_lib.lib().em_csc_to_csc_ex(self._h, ctypes.c_void_p(data.data_ptr()), None, 0.0,
                                          ctypes.c_void_p(out.data_ptr()), prec, flags,
                                          _stream_ptr(stream)))
        return out

    def to_csc_combine(self, dataE: torch.Tensor, dataB: torch.Tensor, k0: float, out=None,
                       stream=None, block_symmetric: bool = False,
                       segment_gather: bool = False) -> torch.Tensor:
        """CSC values of ``dataE - k0**2 * dataB`` in one pass (assembler.py:494)."""
        if dataE.dtype != dataB.dtype or dataE.numel() != dataB.numel():
            raise ValueError("dataE and dataB must match")
        if dataE.numel() != self.nnz_coo or dataE.dtype not in (torch.complex128, torch.complex64):
            raise ValueError(f"COO data must be complex with {self.nnz_coo} entries")
        prec = PREC_FP64 if dataE.dtype == torch.complex128 else PREC_MIXED
        dataE, dataB = dataE.contiguous(), dataB.contiguous()
        if out is None:
            out = torch.empty(self.nnz, dtype=dataE.dtype, device=self.device)
        _check_out(out, self.nnz, dataE.dtype)
        flags = ((_lib.EM_CSC_BLOCK_SYMMETRIC if block_symmetric else 0)
                 | (_lib.EM_CSC_SEGMENT_GATHER if segment_gather else 0))
        check(_lib.lib().em_csc_to_csc_ex(self._h, ctypes.c_void_p(dataE.data_ptr()),
                                          ctypes.c_void_p(dataB.data_ptr()), float(k0),
                                          ctypes.c_void_p(out.data_ptr()), prec, flags,
                                          _stream_ptr(stream)))
        return out

    def to_csc_staged(self, dataE: torch.Tensor, dataB: torch.Tensor | None, k0: float,
                      work: torch.Tensor | None = None, out=None,
                      stream=None) -> torch.Tensor:
        """CSC values of ``dataE - k0**2 * dataB`` (``dataB`` may be None) for blocks that
        are not symmetric (full anisotropic tensors): the contributions are staged
        block-transposed in FP64 in ``work`` (complex128, ``nnz_coo`` entries; allocated
        when None) and summed with contiguous row reads. Equal bit for bit to
        ``to_csc_combine`` / ``to_csc``."""
        if dataE.numel() != self.nnz_coo or dataE.dtype not in (torch.complex128, torch.complex64):
            raise ValueError(f"COO data must be complex with {self.nnz_coo} entries")
        if dataB is not None and (dataB.dtype != dataE.dtype or dataB.numel() != dataE.numel()):
            raise ValueError("dataE and dataB must match")
        prec = PREC_FP64 if dataE.dtype == torch.complex128 else PREC_MIXED
        dataE = dataE.contiguous()
        dataB = dataB.contiguous() if dataB is not None else None
        if work is None:
            work = torch.empty(self.nnz_coo, dtype=torch.complex128, device=self.device)
        _check_out(work, self.nnz_coo, torch.complex128)
        if out is None:
            out = torch.empty(self.nnz, dtype=dataE.dtype, device=self.device)
        _check_out(out, self.nnz, dataE.dtype)
        check(_lib.lib().em_csc_to_csc_staged(
            self._h, ctypes.c_void_p(dataE.data_ptr()),
            None if dataB is None else ctypes.c_void_p(dataB.data_ptr()), float(k0),
            ctypes.c_void_p(work.data_ptr()), ctypes.c_void_p(out.data_ptr()), prec,
            _stream_ptr(stream)))
        return out


def csc_freq_combine(K: torch.Tensor, M: torch.Tensor, k0, out=None,
                     stream=None, C: torch.Tensor | None = None, g=None) -> torch.Tensor:
    """``A = K - k0**2 * M`` for CSC value arrays ``K``, ``M`` of one pattern
    (``em_csc_freq_combine``): ``k0`` a float gives ``[nnz]``, a sequence ``[len(k0), nnz]``
    in one pass over ``K`` and ``M``. complex128 or complex64 (FP64 arithmetic). With
    ``C`` and ``g`` (one value per k0): ``A = K - k0**2 M + 1j * g * C``
    (``em_csc_freq_combine_lossy``, a lossy medium's conductivity term)."""
    if K.dtype != M.dtype or K.numel() != M.numel() or \
            K.dtype not in (torch.complex128, torch.complex64):
        raise ValueError("K and M must be complex CSC value arrays of one pattern")
    if C is not None and (C.dtype != K.dtype or C.numel() != K.numel() or g is None):
        raise ValueError("C must match K and M, with one g per k0")
    K, M = K.contiguous(), M.contiguous()
    prec = PREC_FP64 if K.dtype == torch.complex128 else PREC_MIXED
    scalar = np.isscalar(k0)
    k0a = np.ascontiguousarray(np.atleast_1d(np.asarray(k0, dtype=np.float64)))
    nf, n = int(k0a.shape[0]), K.numel()
    if out is None:
        # complex64 rows start on 16-byte boundaries: an odd length gets a padded row
        pad = 1 if (K.dtype == torch.complex64 and nf > 1 and n % 2) else 0
        out = torch.empty((nf, n + pad), dtype=K.dtype, device=K.device)[:, :n]
    else:
        _check_out(out, nf * n, K.dtype, rows=nf)
    stride = n if out.dim() == 1 else out.stride(0)
    if C is None:
        check(_lib.lib().em_csc_freq_combine(n, ctypes.c_void_p(K.data_ptr()),
                                             ctypes.c_void_p(M.data_ptr()),
                                             k0a.ctypes.data_as(ctypes.c_void_p), nf,
                                             ctypes.c_void_p(out.data_ptr()), stride, prec,
                                             _stream_ptr(stream)))
    else:
        C = C.contiguous()
        ga = np.ascontiguousarray(np.atleast_1d(np.asarray(g, dtype=np.float64)))
        if ga.shape != k0a.shape:
            raise ValueError("one g per k0")
        check(_lib.lib().em_csc_freq_combine_lossy(
            n, ctypes.c_void_p(K.data_ptr()), ctypes.c_void_p(M.data_ptr()),
            ctypes.c_void_p(C.data_ptr()), k0a.ctypes.data_as(ctypes.c_void_p),
            ga.ctypes.data_as(ctypes.c_void_p), nf, ctypes.c_void_p(out.data_ptr()), stride,
            prec, _stream_ptr(stream)))
    if scalar and out.dim() == 2:
        return out[0]
    return out


class FreqCsc:
    """CSC values of a plan's K and M, summed once, and ``A(k0) = K - k0**2 M`` per
    frequency as one streaming pass (``em_csc_freq_combine``).

    The reference converts ``Evec - Bvec * K0**2`` to CSC at every frequency
    (assembler.py:471-494, called per frequency from microwave_3d.py:1464-1470); while the
    element matrices are cached the sums of K and M are too. Each entry is then summed as
    ``sum(E) - k0**2 sum(B)`` instead of ``sum(E - k0**2 B)``: the same matrix up to
    rounding. ``block_symmetric``: the blocks of K and M are symmetric (true of every block
    the element kernels write), which selects the mirrored column gather.
    """

    def __init__(self, cscmap: CscMap, K: torch.Tensor, M: torch.Tensor,
                 block_symmetric: bool = False, stream=None):
        self.map = cscmap
        self.K = cscmap.to_csc(K, block_symmetric=block_symmetric, stream=stream)
        self.M = cscmap.to_csc(M, block_symmetric=block_symmetric, stream=stream)

    @property
    def nnz(self) -> int:
        return self.map.nnz

    def A(self, k0, out=None, stream=None) -> torch.Tensor:
        return csc_freq_combine(self.K, self.M, k0, out=out, stream=stream)


def _check_out(t: torch.Tensor, numel: int, dt, rows: int = 1) -> None:
    """An output of ``numel`` values in ``rows`` rows (one per wavenumber): a contiguous
    1-D tensor, or a 2-D tensor of shape ``(rows, numel // rows)`` whose rows are
    contiguous and do not overlap (row stride at least the row length)."""
    if t.dtype != dt or not t.is_cuda or t.numel() != numel:
        raise ValueError(f"output must be a {dt} CUDA tensor with {numel} elements")
    if t.dim() == 1:
        if not t.is_contiguous():
            raise ValueError("output must be contiguous")
    elif t.dim() == 2:
        n = numel // rows
        if tuple(t.shape) != (rows, n):
            raise ValueError(f"2-D output must have shape ({rows}, {n})")
        if n > 1 and t.stride(1) != 1:
            raise ValueError("output rows must be contiguous")
        if rows > 1 and t.stride(0) < n:
            raise ValueError("output rows must not overlap")
    else:
        raise ValueError("output must be 1-D or 2-D")
