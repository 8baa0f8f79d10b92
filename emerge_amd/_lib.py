"""ctypes binding of libemerge_amd.so (the C-ABI declared in include/emerge_amd.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# EMERGE_AMD_LIB points at another build of the same C-ABI (e.g. one built with different
# compile-time options); the default is the in-tree library.
LIB_PATH = os.environ.get("EMERGE_AMD_LIB", os.path.join(_HERE, "libemerge_amd.so"))

EM_OK = 0
EM_EINVAL = -1
EM_EUNSUPPORTED = -2
EM_EHIP = -3
EM_ENOMEM = -4
EM_EDEGENERATE = -5
EM_ETABLES = -6

EM_MAT_SCALAR = 0
EM_MAT_DIAG = 1
EM_MAT_FULL = 2
EM_MAT_IS_INVERSE = 1

EM_CSC_BLOCK_SYMMETRIC = 1
EM_CSC_SEGMENT_GATHER = 2

EM_TRI_BY_FACE = 0
EM_TRI_COMPACT = 1

EM_INTERP_E = 0
EM_INTERP_CURL = 1

EM_PREC_FP64 = 0
EM_PREC_MIXED = 1
EM_PREC_FP64_REF = 2


class EmError(RuntimeError):
    """A non-zero status returned by the C-ABI, with ``em_last_error()``'s message."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"emerge_amd status {status}: {msg}")
        self.status = status


class DegenerateTetError(EmError):
    pass


class UnsupportedError(EmError):
    pass


class TableError(EmError):
    pass


class em_material_t(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("kind", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("tet_stride", ctypes.c_int64), ("comp_stride", ctypes.c_int64)]


_lib = None

P = ctypes.c_void_p
I64 = ctypes.c_int64
I32 = ctypes.c_int32

_SIGS = {
    "em_last_error": (ctypes.c_char_p, []),
    "em_version": (ctypes.c_char_p, []),
    "em_ned2_plan_create": (ctypes.c_int, [ctypes.POINTER(P), ctypes.c_int, P, I64, P, I64, P, I64,
                                           P, I64, P, P, P, I64, P, I64, P, I32]),
    "em_ned2_plan_destroy": (ctypes.c_int, [P]),
    "em_ned2_plan_set_tuning": (ctypes.c_int, [P, I32]),
    "em_ned2_plan_get_tuning": (ctypes.c_int, [P, ctypes.POINTER(I32)]),
    "em_ned2_plan_set_window": (ctypes.c_int, [P, I64]),
    "em_ned2_plan_get_window": (ctypes.c_int, [P, ctypes.POINTER(I64)]),
    "em_ned2_plan_set_prefetch": (ctypes.c_int, [P, I64]),
    "em_ned2_plan_get_prefetch": (ctypes.c_int, [P, ctypes.POINTER(I64)]),
    "em_ned2_assemble_csc": (ctypes.c_int, [P, P, ctypes.POINTER(em_material_t),
                                            ctypes.POINTER(em_material_t), ctypes.c_double, I32,
                                            P, I32, P]),
    "em_ned2_plan_sizes": (ctypes.c_int, [P, ctypes.POINTER(I64), ctypes.POINTER(I64),
                                          ctypes.POINTER(I64)]),
    "em_ned2_assemble_km": (ctypes.c_int, [P, ctypes.POINTER(em_material_t),
                                           ctypes.POINTER(em_material_t), P, P, I32, P]),
    "em_ned2_assemble_a": (ctypes.c_int, [P, ctypes.POINTER(em_material_t),
                                          ctypes.POINTER(em_material_t), P, I32, P, I64, I32, P]),
    "em_ned2_rowcol": (ctypes.c_int, [P, P, P, P]),
    "em_csc_map_create": (ctypes.c_int, [ctypes.POINTER(P), P, P]),
    "em_csc_map_destroy": (ctypes.c_int, [P]),
    "em_csc_map_sizes": (ctypes.c_int, [P, ctypes.POINTER(I64), ctypes.POINTER(I64),
                                        ctypes.POINTER(I64)]),
    "em_csc_map_export": (ctypes.c_int, [P, P, P, P, P, P, P]),
    "em_csc_to_csc": (ctypes.c_int, [P, P, P, I32, P]),
    "em_csc_to_csc_combine": (ctypes.c_int, [P, P, P, ctypes.c_double, P, I32, P]),
    "em_csc_to_csc_ex": (ctypes.c_int, [P, P, P, ctypes.c_double, P, I32, I32, P]),
    "em_csc_to_csc_staged": (ctypes.c_int, [P, P, P, ctypes.c_double, P, P, I32, P]),
    "em_csc_freq_combine": (ctypes.c_int, [I64, P, P, P, I32, P, I64, I32, P]),
    "em_csc_freq_combine_lossy": (ctypes.c_int, [I64, P, P, P, P, P, I32, P, I64, I32, P]),
    "em_ned2_backward": (ctypes.c_int, [P, P, ctypes.c_double, I32, P, I32, P, P]),
    "em_ned2_pec_mask": (ctypes.c_int, [P, P, I64, P, I64, P, I64, P, P]),
    "em_dof_compact": (ctypes.c_int, [I64, P, P, P, ctypes.POINTER(I64), P]),
    "em_material_prep": (ctypes.c_int, [P, P, P, I64, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, P, P, ctypes.POINTER(I64),
                                        ctypes.POINTER(I32), P]),
    "em_material_prep_noalloc": (ctypes.c_int, [P, P, P, I64, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double, P, P, ctypes.POINTER(I64),
                                                ctypes.POINTER(I32), P]),
    "em_csc_select_count": (ctypes.c_int, [I64, P, P, P, I32, I32, P, I64, P,
                                           ctypes.POINTER(I64), P]),
    "em_csc_select_fill": (ctypes.c_int, [I64, P, P, P, I32, I32, P, P, P, P, P]),
    "em_csc_map_restrict": (ctypes.c_int, [P, P, I64, ctypes.POINTER(P), P]),
    "em_dof_gather": (ctypes.c_int, [P, I64, I64, P, P, I32, P]),
    "em_dof_scatter": (ctypes.c_int, [P, I64, I64, P, P, I32, P]),
    "em_ned2_robin_blocks": (ctypes.c_int, [P, P, I64, P, I64, ctypes.c_double, ctypes.c_double,
                                            P, I32, P]),
    "em_ned2_robin_force": (ctypes.c_int, [P, P, I64, P, P, I64, P, P, P, P, P]),
    "em_tri_rowcol": (ctypes.c_int, [P, I64, P, I64, P, P, P]),
    "em_csc_map_find": (ctypes.c_int, [P, P, P, I64, P, P, P]),
    "em_csc_add": (ctypes.c_int, [P, I32, P, P, I64, ctypes.c_double, ctypes.c_double, P]),
    "em_ned2_locate": (ctypes.c_int, [P, P, I64, P, I64, P, P]),
    "em_ned2_interp": (ctypes.c_int, [P, P, I64, P, P, I32, P, P, P]),
    "em_heat_plan_create": (ctypes.c_int, [ctypes.POINTER(P), ctypes.c_int, P, I64, P, I64, P,
                                           I64]),
    "em_heat_plan_destroy": (ctypes.c_int, [P]),
    "em_heat_assemble": (ctypes.c_int, [P, P, I32, I64, I64, P, I64, P, P, P]),
    "em_heat_rowcol": (ctypes.c_int, [P, P, P, P]),
    "em_csc_map_create_heat": (ctypes.c_int, [ctypes.POINTER(P), P, P]),
    "em_csc_to_csc_real": (ctypes.c_int, [P, P, P, ctypes.c_double, P, ctypes.c_int, P]),
    "em_graph_begin_capture": (ctypes.c_int, [P]),
    "em_graph_end_capture": (ctypes.c_int, [P, ctypes.POINTER(P)]),
    "em_graph_launch": (ctypes.c_int, [P, P]),
    "em_graph_destroy": (ctypes.c_int, [P]),
}

# every symbol include/emerge_amd.h declares
EXPORTED = tuple(_SIGS)


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m emerge_amd.build`")
    # torch ships its own libamdhip64; loading it first makes the library bind to that
    # runtime (same soname), so torch's device pointers and streams are valid in it. The
    # other order leaves two HIP runtimes in the process, and the second sees no device.
    import torch  # noqa: F401
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status: int) -> None:
    if status == EM_OK:
        return
    msg = lib().em_last_error().decode(errors="replace")
    cls = {EM_EDEGENERATE: DegenerateTetError, EM_EUNSUPPORTED: UnsupportedError,
           EM_ETABLES: TableError}.get(status, EmError)
    raise cls(status, msg)
