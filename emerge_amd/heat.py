"""Heat-conduction assembly with second-order Lagrange elements, on the GPU.

Mirrors src/emerge/_emerge/physics/heatconduction/assembly/grad.py:

* ``tet_stiffness_matrix(field, cond_termal)`` (:417-437) -- conductance COO values and
  the ``CSCMapping`` of the pattern;
* ``tet_mass_matrix(field, rho_cp)`` (:440-462) -- capacity COO values and the mapping.

``HeatPlan`` / ``HeatCscMap`` are the device-resident forms (``em_heat_*``,
``em_csc_map_create_heat``, ``em_csc_to_csc_real``); the returned mapping's ``to_csc``
gives a SciPy ``csc_matrix`` like the reference's ``CSCMapping.to_csc`` for float data.
"""
from __future__ import annotations

import ctypes
import weakref

import numpy as np
import torch

from . import _lib
from ._lib import check
from .assembly import _stream_ptr, to_host


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("emerge_amd needs a HIP device")
    return torch.device("cuda", torch.cuda.current_device())


class HeatPlan:
    """Device copy of a Lagrange-2 field's mesh and DOF map."""

    def __init__(self, nodes, tets, tet_to_field, n_field, device=None):
        self.device = device or _device()
        nodes = np.ascontiguousarray(nodes, np.float64)
        tets = np.ascontiguousarray(tets, np.int64)
        t2f = np.ascontiguousarray(tet_to_field, np.int64)
        if tets.ndim != 2 or tets.shape[0] != 4 or t2f.shape != (10, tets.shape[1]):
            raise ValueError("tets must be (4, T) and tet_to_field (10, T)")
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(_lib.lib().em_heat_plan_create(ctypes.byref(h), self.device.index or 0,
                                                 nodes.ctypes.data_as(ctypes.c_void_p),
                                                 nodes.shape[1],
                                                 tets.ctypes.data_as(ctypes.c_void_p),
                                                 tets.shape[1],
                                                 t2f.ctypes.data_as(ctypes.c_void_p),
                                                 int(n_field)))
        self._h = h
        self.n_tets = tets.shape[1]
        self.n_field = int(n_field)
        self.nnz_coo = 100 * self.n_tets

    @classmethod
    def from_field(cls, field, device=None) -> "HeatPlan":
        return cls(field.mesh.nodes, field.mesh.tets, field.tet_to_field, field.n_field, device)

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().em_heat_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _cond(self, cond):
        """Device float64 view + (kind, tet stride, comp stride) of a conductivity."""
        t = torch.as_tensor(cond) if not isinstance(cond, torch.Tensor) else cond
        t = t.to(device=self.device, dtype=torch.float64)
        T = self.n_tets
        if t.dim() == 0:
            return t.reshape(1).contiguous(), _lib.EM_MAT_SCALAR, 0, 0
        if t.shape == (T,):
            return t.contiguous(), _lib.EM_MAT_SCALAR, 1, 0
        if t.shape == (T, 3):
            t = t.contiguous()
            return t, _lib.EM_MAT_DIAG, 3, 1
        if t.shape == (T, 3, 3):
            t = t.contiguous()
            return t, _lib.EM_MAT_FULL, 9, 1
        if t.shape == (3, 3, T):  # the reference's layout
            t = t.contiguous()
            return t, _lib.EM_MAT_FULL, 1, T
        raise ValueError(f"conductivity shape {tuple(t.shape)} not understood")

    def assemble(self, cond=None, rho_cp=None, stream=None):
        """(K, M) float64 device COO values (None where not asked for)."""
        K = M = None
        c = k = ts = cs = None
        if cond is not None:
            c, k, ts, cs = self._cond(cond)
            K = torch.empty(self.nnz_coo, dtype=torch.float64, device=self.device)
        r = None
        rts = 0
        if rho_cp is not None:
            r = torch.as_tensor(rho_cp).to(device=self.device, dtype=torch.float64).reshape(-1)
            if r.numel() not in (1, self.n_tets):
                raise ValueError("rho_cp must be a scalar or one value per tet")
            rts = 0 if r.numel() == 1 else 1
            r = r.contiguous()
            M = torch.empty(self.nnz_coo, dtype=torch.float64, device=self.device)
        check(_lib.lib().em_heat_assemble(self._h, _vp(c), 0 if k is None else k,
                                          0 if ts is None else ts, 0 if cs is None else cs,
                                          _vp(r), rts, _vp(K), _vp(M), _stream_ptr(stream)))
        return K, M

    def rowcol(self, stream=None):
        rows = torch.empty(max(self.nnz_coo, 1), dtype=torch.int64, device=self.device)
        cols = torch.empty_like(rows)
        check(_lib.lib().em_heat_rowcol(self._h, _vp(rows), _vp(cols), _stream_ptr(stream)))
        return rows[:self.nnz_coo], cols[:self.nnz_coo]


class HeatCscMap:
    """Device CSC pattern of a heat plan (``CSCMapping.from_rowcol``)."""

    def __init__(self, plan: HeatPlan, stream=None):
        h = ctypes.c_void_p()
        check(_lib.lib().em_csc_map_create_heat(ctypes.byref(h), plan._h, _stream_ptr(stream)))
        self._h = h
        self.device = plan.device
        n, nnz, nnz_coo = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(_lib.lib().em_csc_map_sizes(h, ctypes.byref(n), ctypes.byref(nnz),
                                          ctypes.byref(nnz_coo)))
        self.N, self.nnz, self.nnz_coo = n.value, nnz.value, nnz_coo.value

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().em_csc_map_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def export(self, *names, stream=None) -> dict:
        sizes = {"indptr": self.N + 1, "indices": self.nnz, "perm": self.nnz_coo,
                 "target_csc": self.nnz_coo, "col_offsets": self.N + 1}
        out, ptrs = {}, []
        for k in ("indptr", "indices", "perm", "target_csc", "col_offsets"):
            if k in names:
                out[k] = torch.empty(sizes[k], dtype=torch.int64, device=self.device)
                ptrs.append(_vp(out[k]))
            else:
                ptrs.append(None)
        check(_lib.lib().em_csc_map_export(self._h, *ptrs, _stream_ptr(stream)))
        return out

    def to_csc(self, data: torch.Tensor, data_b: torch.Tensor | None = None, scale: float = 0.0,
               out=None, stream=None, segment_gather: bool = False,
               block_symmetric: bool = False) -> torch.Tensor:
        """CSC values of ``data`` (+ ``scale * data_b``), float64 on the device.
        ``segment_gather`` selects the per-entry gather instead of the column-incidence one;
        ``block_symmetric`` (every 10x10 block bit-symmetric: capacity blocks, conductance
        blocks of a scalar or diagonal conductivity) reads each entry from its mirror. All
        give the same result bit for bit."""
        if data.dtype != torch.float64 or data.numel() != self.nnz_coo:
            raise ValueError(f"data must be float64 with {self.nnz_coo} entries")
        data = data.contiguous()
        if data_b is not None:
            if data_b.dtype != torch.float64 or data_b.numel() != self.nnz_coo:
                raise ValueError("data_b must match data")
            data_b = data_b.contiguous()
        if out is None:
            out = torch.empty(self.nnz, dtype=torch.float64, device=self.device)
        flags = (_lib.EM_CSC_SEGMENT_GATHER if segment_gather else 0) | \
            (_lib.EM_CSC_BLOCK_SYMMETRIC if block_symmetric else 0)
        check(_lib.lib().em_csc_to_csc_real(self._h, _vp(data), _vp(data_b), float(scale),
                                            _vp(out), flags, _stream_ptr(stream)))
        return out


class HeatCSCMapping:
    """Host-facing mapping (the reference's ``CSCMapping`` for float data)."""

    def __init__(self, dev: HeatCscMap):
        self._dev = dev
        ex = dev.export("indptr", "indices")
        self.indptr = to_host(ex["indptr"])
        self.indices = to_host(ex["indices"])
        self.N, self.nnz, self.nnz_coo = dev.N, dev.nnz, dev.nnz_coo

    def to_csc(self, data):
        import scipy.sparse as sp
        t = torch.as_tensor(np.ascontiguousarray(data, np.float64)).to(self._dev.device)
        vals = to_host(self._dev.to_csc(t))
        return sp.csc_matrix((vals, self.indices, self.indptr), shape=(self.N, self.N))


_PLANS: "weakref.WeakKeyDictionary" = weakref.WeakKeyDictionary()


def _plan_for(field) -> HeatPlan:
    p = _PLANS.get(field)
    if p is None:
        p = HeatPlan.from_field(field)
        _PLANS[field] = p
    return p


def tet_stiffness_matrix(field, cond_termal):
    """grad.py:417-437: (conductance COO values float64 [100 T], CSCMapping)."""
    plan = _plan_for(field)
    K, _ = plan.assemble(cond=np.asarray(cond_termal, np.float64))
    return to_host(K), HeatCSCMapping(HeatCscMap(plan))


def tet_mass_matrix(field, rho_cp):
    """grad.py:440-462: (capacity COO values float64 [100 T], CSCMapping)."""
    plan = _plan_for(field)
    _, M = plan.assemble(rho_cp=np.asarray(rho_cp, np.float64))
    return to_host(M), HeatCSCMapping(HeatCscMap(plan))
