"""Drop-in for the reference's curl-curl assembly entry point.

``tet_mass_stiffness_matrices(field, er, ur, conductor_tets, cscmap=None)`` has the
signature, argument meaning and return types of
``src/emerge/_emerge/physics/microwave/assembly/curlcurl.py:222-270``:

* ``field`` is a Nedelec-2 field object (``mesh.tets/tris/edges/nodes/tet_to_edge/
  tet_to_tri``, ``get_tet_to_field()``, ``n_field``, ``dofcodes3d``);
* ``er``/``ur`` are complex ``(3, 3, T)`` tensors (``ur`` is inverted per tet, as the
  reference's ``matinv`` does);
* ``conductor_tets`` lists tets left out of the assembly (``curlcurl.py:248-250``);
* it returns ``(dataE, dataB, cscmap)``: NumPy complex128 COO values of the stiffness
  and mass blocks, and a ``CSCMapping`` whose ``to_csc(data)`` returns a SciPy
  ``csc_matrix`` (``csc_cast.py:55-63``).

The work runs on the GPU: the HIP kernels compute the element blocks and the CSC map is
built and applied on the device. Host arrays cross PCIe in both directions, since the
reference's callers hold NumPy arrays; the returned arrays live in pinned host memory
(a copy into pageable memory runs at a few GB/s and would make the call slower than the
CPU reference). The device plan of a field is cached.
"""
from __future__ import annotations

import weakref

import numpy as np
import torch

from .assembly import CscMap, Ned2Plan, to_host

_PLANS: "weakref.WeakKeyDictionary" = weakref.WeakKeyDictionary()


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("emerge_amd assembly needs a GPU")
    return torch.device("cuda", torch.cuda.current_device())


def plan_for(field, conductor_tets=None) -> Ned2Plan:
    """The cached device plan of ``field`` for this conductor set."""
    T = field.mesh.tets.shape[1]
    conductor_tets = np.asarray([] if conductor_tets is None else conductor_tets)
    key = conductor_tets.astype(np.int64).tobytes()
    entry = _PLANS.get(field)
    if entry is not None and entry[0] == key:
        return entry[1]
    tetids = None
    if conductor_tets.shape[0] > 0:
        tetids = np.delete(np.arange(T), conductor_tets.astype(np.int64))
    plan = Ned2Plan.from_field(field, tetids=tetids, device=_device())
    _PLANS[field] = (key, plan)
    return plan


class CSCMapping:
    """GPU-built counterpart of the reference ``CSCMapping`` (csc_cast.py:29-63).

    ``indptr``/``indices`` are host int64 arrays as in the reference; ``to_csc(data)``
    takes host COO values and returns a SciPy ``csc_matrix`` with sorted indices and
    duplicates summed. ``to_csc_device`` keeps everything on the GPU.
    """

    def __init__(self, dev: CscMap):
        self._dev = dev
        arrs = dev.export("indptr", "indices")
        self.indptr = to_host(arrs["indptr"])
        self.indices = to_host(arrs["indices"])
        self.N = dev.N
        self.nnz = dev.nnz
        self.nnz_coo = dev.nnz_coo

    @classmethod
    def from_plan(cls, plan: Ned2Plan) -> "CSCMapping":
        return cls(CscMap(plan))

    def _lazy(self, name):
        return to_host(self._dev.export(name)[name])

    @property
    def perm(self) -> np.ndarray:
        return self._lazy("perm")

    @property
    def target_csc(self) -> np.ndarray:
        return self._lazy("target_csc")

    @property
    def col_offsets(self) -> np.ndarray:
        return self._lazy("col_offsets")

    def to_csc_device(self, data: torch.Tensor) -> torch.Tensor:
        """CSC values of device COO values. Blocks that are symmetric (the element
        matrices of any tensors, and elementwise combinations of them such as
        E - K0**2 B) are summed by the mirrored column gather, others by the staged one;
        both give the plain gather's result bit for bit (up to the sign of zero sums)."""
        data = data.contiguous()
        if data.numel() == 0:
            return self._dev.to_csc(data)
        blocks = data.view(-1, 20, 20)
        if bool(torch.equal(blocks, blocks.transpose(1, 2))):
            return self._dev.to_csc(data, block_symmetric=True)
        return self._dev.to_csc_staged(data, None, 0.0)

    def to_csc(self, data):
        from scipy.sparse import csc_matrix
        if isinstance(data, torch.Tensor):
            vals = to_host(self.to_csc_device(data.to(self._dev.device)))
        else:
            arr = np.asarray(data)
            if np.iscomplexobj(arr):
                t = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.complex128))
                vals = to_host(self.to_csc_device(t.to(self._dev.device)))
            else:
                # real data: the imaginary part of the complex sum is zero
                t = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float64)).to(torch.complex128)
                vals = to_host(self.to_csc_device(t.to(self._dev.device)).real.contiguous())
        return csc_matrix((vals, self.indices, self.indptr), shape=(self.N, self.N))


def tet_mass_stiffness_matrices(field, er: np.ndarray, ur: np.ndarray,
                                conductor_tets: np.ndarray, cscmap: CSCMapping | None = None,
                                precision: str = "fp64"):
    """GPU drop-in for ``tet_mass_stiffness_matrices`` (curlcurl.py:222-270).

    Returns ``(dataE, dataB, cscmap)`` with ``dataE`` the stiffness (``mu_r^-1``) and
    ``dataB`` the mass (``eps_r``) COO values, complex128 NumPy arrays of length
    ``400 * (T - len(conductor_tets))`` in the reference's layout.

    ``precision`` (not in the reference's signature; the default keeps its calls
    unchanged): ``"fp64"`` gives the exact element integrals; ``"fp64_ref"`` repeats the
    reference's own 11-point arithmetic in absolute coordinates, whose cofactor sums
    cancel on thin tets (the two differ by up to ~1e-7 relative on seeded Delaunay
    meshes, ~1e-15 on well-shaped meshes near the origin).
    """
    if precision not in ("fp64", "fp64_ref"):
        raise ValueError("the drop-in returns complex128: precision 'fp64' or 'fp64_ref'")
    plan = plan_for(field, conductor_tets)
    K, M = plan.assemble_km(er, ur, precision)
    if cscmap is None:
        cscmap = CSCMapping.from_plan(plan)
    return to_host(K), to_host(M), cscmap
