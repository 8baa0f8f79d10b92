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
        dt = torch.complex128 if precision in ("fp64", 0) else torch.complex64
        if out is None:
            out = torch.empty((self.k0s.shape[0], plan.nnz_coo), dtype=dt, device=plan.device)
        self.out = out
        self.precision = precision
        self.graph = Graph(lambda: plan.assemble_a(self.er, self.ur, self.k0s, precision,
                                                   out=self.out))

    def replay(self, stream=None) -> torch.Tensor:
        self.graph.replay(stream)
        return self.out
