"""Material tensors of one frequency on the device (``em_material_prep``).

Mirrors the preparation in ``Assembler.assemble_freq_matrix``
(src/emerge/_emerge/physics/microwave/assembly/assembler.py:439-469; the same in
``assemble_scattering_matrix`` :819-868 and ``assemble_eig_matrix`` :1119-1142) that
precedes ``tet_mass_stiffness_matrices``:

* ``er = er * (1 - 1j*tand) - 1j*cond / (W0*EPS0)`` with ``W0 = 2*pi*f`` (:423, :453);
* ``is_frequency_dependent``: some conductivity component with ``0 < cond < mw_3d_peclim``
  (:455-457), which tells the caller not to reuse cached matrices;
* ``conductor_tets``: the tets whose ``cond[0,0]`` exceeds ``mw_3d_peclim`` (1e8) or
  ``mw_3d_surfimplim`` (1e4) (:461-469, settings.py:35-38), ascending; they are left out of
  the assembly and their DOFs become PEC.

The reference builds the conductor list with a Python loop over the tets; here one kernel
does the arithmetic, the comparisons and the compaction.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check
from .assembly import _stream_ptr

PECLIM = 1e8       # settings.py:35
SURFIMPLIM = 1e4   # settings.py:38


def _dev_tensor(x, device, T=None):
    t = x if isinstance(x, torch.Tensor) else torch.from_numpy(np.asarray(x))
    t = t.to(device=device, dtype=torch.complex128).contiguous()
    if t.dim() != 3 or t.shape[:2] != (3, 3) or (T is not None and t.shape[2] != T):
        raise ValueError(f"material tensors must be (3, 3, T); got {tuple(t.shape)}")
    return t


def prepare_materials(er, tand, cond, frequency: float, peclim: float = PECLIM,
                      surfimplim: float = SURFIMPLIM, device=None, out=None, stream=None):
    """Returns ``(er, conductor_tets, is_frequency_dependent)``: ``er`` a device complex128
    (3,3,T) tensor (``out`` if given; ``out`` may be the input ``er`` tensor itself),
    ``conductor_tets`` an int64 NumPy array, ascending, as the reference passes to
    ``tet_mass_stiffness_matrices`` and to the PEC DOF collection."""
    if device is None:
        device = er.device if isinstance(er, torch.Tensor) and er.is_cuda else \
            torch.device("cuda", torch.cuda.current_device())
    e = _dev_tensor(er, device)
    T = e.shape[2]
    t = _dev_tensor(tand, device, T)
    c = _dev_tensor(cond, device, T)
    if out is None:
        out = torch.empty_like(e)
    elif not isinstance(out, torch.Tensor) or out.shape != e.shape \
            or out.dtype != torch.complex128 or not out.is_contiguous() or out.device != e.device:
        raise ValueError("out must be a contiguous complex128 (3, 3, T) tensor on the device")
    ids = torch.empty(max(T, 1), dtype=torch.int64, device=device)
    n = ctypes.c_int64()
    fd = ctypes.c_int32()
    check(_lib.lib().em_material_prep(e.data_ptr(), t.data_ptr(), c.data_ptr(), T,
                                      2 * np.pi * frequency, peclim, surfimplim,
                                      out.data_ptr(), ids.data_ptr(), ctypes.byref(n),
                                      ctypes.byref(fd), _stream_ptr(stream)))
    return out, ids[:n.value].cpu().numpy(), bool(fd.value)
