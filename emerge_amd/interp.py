"""Point location and Nedelec-2 field interpolation on the GPU.

Mirrors the reference's interpolation path:

* ``interpolate_index(field, xs, ys, zs, tetids)`` -- ``Nedelec2.interpolate_index``
  (elements/nedelec2.py:319-340) over ``index_interp`` (compiled/base/index_inter.py:
  262-298): the tet of every point (-1 when none);
* ``ned2_tet_interp`` / ``ned2_tet_interp_curl`` -- compiled/base/interp.py:286-387 and
  :410-519 for points with known tets;
* ``interpolate_Ef(field, solution, basis, origin, tetids)`` -- FemData.interpolate_Ef
  (elements/femdata.py:125-162): a callable ``f(xs, ys, zs) -> array([Ex, Ey, Ez])``.

The kernels are ``em_ned2_locate`` and ``em_ned2_interp`` (csrc/interp.hip).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check
from .assembly import Ned2Plan, _stream_ptr


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _plan(field, plan):
    if plan is not None:
        return plan
    from .curlcurl import plan_for
    return plan_for(field)


def _coords(plan, xs, ys, zs):
    c = np.ascontiguousarray(np.stack([np.asarray(xs, np.float64).ravel(),
                                       np.asarray(ys, np.float64).ravel(),
                                       np.asarray(zs, np.float64).ravel()]))
    return torch.as_tensor(c, device=plan.device)


def locate(plan: Ned2Plan, coords: torch.Tensor, tetids=None, stream=None) -> torch.Tensor:
    """Tet of every point of ``coords`` (device float64 (3, n)) among ``tetids`` (all plan
    tets when None): the largest containing id, -1 when none."""
    coords = coords.to(torch.float64).contiguous()
    n = coords.shape[1]
    out = torch.empty(max(n, 1), dtype=torch.int64, device=plan.device)
    ids = None
    if tetids is not None:
        ids = torch.as_tensor(np.asarray(tetids, np.int64) if not isinstance(tetids, torch.Tensor)
                              else tetids, device=plan.device).to(torch.int64).contiguous()
    check(_lib.lib().em_ned2_locate(plan._h, _vp(coords), n, _vp(ids),
                                    0 if ids is None else ids.numel(), _vp(out),
                                    _stream_ptr(stream)))
    return out[:n]


def interp(plan: Ned2Plan, coords: torch.Tensor, solution: torch.Tensor, tet_of: torch.Tensor,
           curl: bool = False, c_tet: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """(3, n) complex128: the field (or ``c_tet * curl``) of ``solution`` at the points."""
    coords = coords.to(torch.float64).contiguous()
    n = coords.shape[1]
    sol = solution.to(device=plan.device, dtype=torch.complex128).contiguous()
    if sol.numel() != plan.n_field:
        raise ValueError(f"solution must have {plan.n_field} entries")
    tet_of = tet_of.to(device=plan.device, dtype=torch.int64).contiguous()
    ct = None
    if c_tet is not None:
        ct = torch.as_tensor(c_tet).to(device=plan.device, dtype=torch.complex128).contiguous()
        if ct.numel() != plan.n_tets:
            raise ValueError("c_tet needs one value per tet")
    out = torch.empty((3, max(n, 1)), dtype=torch.complex128, device=plan.device)
    which = _lib.EM_INTERP_CURL if curl else _lib.EM_INTERP_E
    check(_lib.lib().em_ned2_interp(plan._h, _vp(coords), n, _vp(tet_of), _vp(sol), which,
                                    _vp(ct), _vp(out), _stream_ptr(stream)))
    return out[:, :n]


def interpolate_index(field, xs, ys, zs, tetids=None, plan=None) -> np.ndarray:
    """nedelec2.py:319-340: tet of each point (-1 when none)."""
    plan = _plan(field, plan)
    return locate(plan, _coords(plan, xs, ys, zs), tetids).cpu().numpy()


def ned2_tet_interp(field, coords, solution, tet_mapping, plan=None):
    """interp.py:286-387: (Ex, Ey, Ez) at coords (3, n) for points in tets tet_mapping."""
    plan = _plan(field, plan)
    c = torch.as_tensor(np.ascontiguousarray(coords, np.float64), device=plan.device)
    E = interp(plan, c, torch.as_tensor(np.asarray(solution)),
               torch.as_tensor(np.asarray(tet_mapping, np.int64)))
    return tuple(E.cpu().numpy())


def ned2_tet_interp_curl(field, coords, solution, c, tet_mapping, plan=None):
    """interp.py:410-519: c[tet] * curl E at the points."""
    plan = _plan(field, plan)
    cc = torch.as_tensor(np.ascontiguousarray(coords, np.float64), device=plan.device)
    E = interp(plan, cc, torch.as_tensor(np.asarray(solution)),
               torch.as_tensor(np.asarray(tet_mapping, np.int64)), curl=True,
               c_tet=torch.as_tensor(np.asarray(c, np.complex128)))
    return tuple(E.cpu().numpy())


def interpolate_Ef(field, solution, basis=None, origin=None, tetids=None, plan=None):
    """femdata.py:125-162: f(xs, ys, zs) -> array([Ex, Ey, Ez]) in the given basis."""
    plan = _plan(field, plan)
    basis = np.eye(3) if basis is None else np.asarray(basis, np.float64)
    origin = np.zeros(3) if origin is None else np.asarray(origin, np.float64)
    ibasis = np.linalg.pinv(basis)
    sol = torch.as_tensor(np.asarray(solution, np.complex128), device=plan.device)

    def func(xs, ys, zs):
        xyz = np.array([np.ravel(xs), np.ravel(ys), np.ravel(zs)], dtype=np.float64)
        xyzg = basis @ (xyz + origin[:, None])
        c = torch.as_tensor(np.ascontiguousarray(xyzg), device=plan.device)
        tet = locate(plan, c, tetids)
        E = interp(plan, c, sol, tet).cpu().numpy()
        return ibasis @ E

    return func
