"""em_material_prep_noalloc (include/emerge_amd.h) through the C-ABI against
em_material_prep on the same inputs: the prepared tensor bit for bit, the conductor ids,
their count and the frequency-dependence flag, on two calls in a row, the second smaller
(the fixed scratch keeps the first call's contents) and with PEC-only conductors (a flag
left over from the first call would show)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F = 9.7e9


def _inputs(T, seed, finite_conductivity):
    import torch
    rng = np.random.default_rng(seed)
    er = rng.uniform(1, 10, (3, 3, T)) + 1j * rng.normal(size=(3, 3, T))
    tand = rng.uniform(0, 0.02, (3, 3, T)) + 1j * rng.normal(size=(3, 3, T)) * 1e-3
    cond = np.zeros((3, 3, T), complex)
    pick = rng.choice(T, max(1, T // 50), replace=False)
    cond[0, 0, pick] = 2e8  # conductors (above mw_3d_peclim)
    if finite_conductivity:  # components strictly between 0 and mw_3d_peclim
        cond[1, 1, pick[::2]] = 1e4 + 1j
        cond[2, 2, rng.choice(T, 5, replace=False)] = 3.5
    return [torch.as_tensor(np.ascontiguousarray(a), device="cuda") for a in (er, tand, cond)]


def _call(fn, er, tand, cond):
    import torch
    T = er.shape[2]
    out = torch.full_like(er, complex("nan"))
    ids = torch.full((T,), -7, dtype=torch.int64, device="cuda")
    n = ctypes.c_int64(-1)
    fd = ctypes.c_int32(-1)
    st = fn(ctypes.c_void_p(er.data_ptr()), ctypes.c_void_p(tand.data_ptr()),
            ctypes.c_void_p(cond.data_ptr()), T, 2 * np.pi * F, 1e8, 1e4,
            ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ids.data_ptr()), ctypes.byref(n),
            ctypes.byref(fd), None)
    torch.cuda.synchronize()
    return st, out, ids[:max(n.value, 0)].cpu().numpy(), n.value, fd.value


def test_noalloc_matches_material_prep_on_two_calls():
    import torch
    from emerge_amd import _lib
    L = _lib.lib()
    for T, seed, finite in ((600_001, 1, True), (4_099, 2, False)):
        er, tand, cond = _inputs(T, seed, finite)
        st_a, out_a, ids_a, n_a, fd_a = _call(L.em_material_prep, er, tand, cond)
        st_b, out_b, ids_b, n_b, fd_b = _call(L.em_material_prep_noalloc, er, tand, cond)
        assert st_a == 0 and st_b == 0, L.em_last_error()
        assert torch.equal(out_a, out_b)
        assert n_b == n_a == ids_a.size and n_a >= T // 50
        assert np.array_equal(ids_b, ids_a) and np.all(np.diff(ids_b) > 0)
        assert fd_b == fd_a == int(finite)
