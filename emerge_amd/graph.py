"""HIP graphs of assembly launches (``em_graph_*`` in include/emerge_amd.h).

A frequency sweep (one ``MODE_AF`` launch per 128 wavenumbers) or a multi-mesh batch
is recorded once with stream capture and replayed with a single graph launch, so the
per-call host work (argument marshalling, material descriptors, launch overhead) is paid
once. Captured callables must use preallocated outputs and prebuilt materials.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import check


class Graph:
    """Record the launches ``fn()`` issues on a private stream; ``replay()`` runs them."""

    def __init__(self, fn, stream: torch.cuda.Stream | None = None):
        L = _lib.lib()
        self.stream = stream if stream is not None else torch.cuda.Stream()
        h = ctypes.c_void_p()
        sp = ctypes.c_void_p(self.stream.cuda_stream)
        torch.cuda.current_stream().synchronize()
        with torch.cuda.stream(self.stream):
            check(L.em_graph_begin_capture(sp))
            try:
                fn()
            finally:
                st = L.em_graph_end_capture(sp, ctypes.byref(h))
            check(st)
        self._h = h

    def replay(self, stream: torch.cuda.Stream | None = None) -> None:
        s = stream if stream is not None else torch.cuda.current_stream()
        check(_lib.lib().em_graph_launch(self._h, ctypes.c_void_p(s.cuda_stream)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().em_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SweepGraph:
    """A whole frequency sweep A_f = K - k0_f^2 M (f = 0..F-1) as one graph.

    ``out`` is ``[F, nnz_coo]`` (complex128 for ``"fp64"``, complex64 for ``"mixed"``);
    ``replay()`` re-assembles every frequency in one graph launch.
    """

    def __init__(self, plan, er, ur, k0s, precision="mixed", out=None):
        import numpy as np
        self.plan = plan
        self.er = plan.material(er)
        self.ur = plan.material(ur)
        self.k0s = np.ascontiguousarray(np.asarray(k0s, np.float64).ravel())
        from .assembly import _out_dtype, _prec
        dt = _out_dtype(_prec(precision))
        if out is None:
            out = torch.empty((self.k0s.shape[0], plan.nnz_coo), dtype=dt, device=plan.device)
        self.out = out
        self.precision = precision
        self.graph = Graph(lambda: plan.assemble_a(self.er, self.ur, self.k0s, precision,
                                                   out=self.out))

    def replay(self, stream=None) -> torch.Tensor:
        self.graph.replay(stream)
        return self.out


class CscSweepGraph:
    """A frequency sweep in CSC form as one graph: the element matrices K and M of every
    tet (``em_ned2_assemble_km``), their CSC sums (two column gathers) and every
    ``A_f = K - k0_f**2 M`` (``em_csc_freq_combine``, one pass per 128 wavenumbers).

    This is the reference's sweep (microwave_3d.py:1464-1470 calling
    ``assemble_freq_matrix``; ``cscmap.to_csc(Evec - Bvec * K0**2)``, assembler.py:494,
    at each frequency) with K and M shared. ``out`` is ``[F, nnz]`` CSC values (pattern:
    ``cscmap.export("indptr", "indices")``); complex128 for ``"fp64"``, complex64 for
    ``"mixed"``. The COO blocks and the CSC sums of K and M stay allocated with the graph.
    """

    def __init__(self, plan, cscmap, er, ur, k0s, precision="mixed", out=None):
        import numpy as np
        from .assembly import _out_dtype, _prec, csc_freq_combine
        self.plan = plan
        self.map = cscmap
        self.er = plan.material(er)
        self.ur = plan.material(ur)
        self.k0s = np.ascontiguousarray(np.asarray(k0s, np.float64).ravel())
        prec = _prec(precision)
        if prec == _lib.EM_PREC_FP64_REF:
            raise ValueError("the CSC sweep takes 'fp64' or 'mixed'")
        dt = _out_dtype(prec)
        dev = plan.device
        self.K = torch.empty(plan.nnz_coo, dtype=dt, device=dev)
        self.M = torch.empty(plan.nnz_coo, dtype=dt, device=dev)
        self.K_csc = torch.empty(cscmap.nnz, dtype=dt, device=dev)
        self.M_csc = torch.empty(cscmap.nnz, dtype=dt, device=dev)
        if out is None:
            out = torch.empty((self.k0s.shape[0], cscmap.nnz), dtype=dt, device=dev)
        self.out = out

        def record():
            plan.assemble_km(self.er, self.ur, precision, out=(self.K, self.M))
            # the element blocks are symmetric whatever the tensors (i <= j mirrored, as
            # curlcurl.py:357-371): mirrored-read gathers
            cscmap.to_csc(self.K, out=self.K_csc, block_symmetric=True)
            cscmap.to_csc(self.M, out=self.M_csc, block_symmetric=True)
            csc_freq_combine(self.K_csc, self.M_csc, self.k0s, out=self.out)

        self.graph = Graph(record)

    def replay(self, stream=None) -> torch.Tensor:
        self.graph.replay(stream)
        return self.out
