"""Autotuner for the assembly kernels' launch shape.

Three knobs, tuned together:

* tets per workgroup tile (32, 64 or 128): the latency of the per-tile geometry phase and
  the tail of partly filled tiles against occupancy;
* tets per window of the tile order (512 .. 16,384): each window's tets are grouped by
  vertex order and its tiles run on one XCD. Small windows keep the blocks written at the
  same time close together in memory (the kernel is bound by its store stream); windows so
  small that a vertex-order group holds a few tets leave the tiles nearly empty. The best
  window moves with the mesh size and the other knobs (512-704 on a 2.8M-tet mesh; on an
  8M one 1,024-4,096 as one launch, 640-704 prefetched);
* prefetched launches (0 or chunks of 2^20 tets): each chunk of the tile list launched
  right after a kernel that reads its inputs into the memory-side cache. It pays where
  the inputs are not already there (a mesh whose inputs exceed the 256 MB cache, or
  launches after other work) and costs the extra reads where they are.

The best choice depends on the precision, the output mode (one k0 or a sweep), the material
kinds and the mesh size. ``tune`` times each candidate pair with HIP events on the launch
stream, applies the fastest to the plan and remembers it in a JSON cache keyed by all of
the above, so later plans of the same kind skip the timing.

Cache file: ``$EMERGE_AMD_TUNE_CACHE`` or ``~/.cache/emerge_amd/autotune.json``.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np
import torch

CANDIDATES = (32, 64, 128)
WINDOWS = (512, 640, 768, 1024, 4096, 16384)
PREFETCH = (0, 1 << 20)


def cache_path() -> str:
    return os.environ.get("EMERGE_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "emerge_amd",
                                       "autotune.json"))


def _load() -> dict:
    try:
        with open(cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _save(d: dict) -> None:
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + f".tmp{os.getpid()}"
        with open(tmp, "w") as fh:
            json.dump(d, fh, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass  # a read-only home only disables persistence


def tuning_key(plan, er, ur, nf: int, precision: str) -> str:
    props = torch.cuda.get_device_properties(plan.device)
    arch = getattr(props, "gcnArchName", props.name)
    size = int(round(math.log2(max(plan.n_assy, 1))))
    mode = "A" if nf == 1 else f"AF{min(nf, 128)}"
    return f"{arch}|{precision}|{mode}|eps{er.kind}|mu{ur.kind}|T2^{size}"


def _time_cold(fn, reps: int, junk) -> float:
    """Median of single launches, each after a fill of ``junk`` on the same stream (which
    evicts the launch's inputs from the memory-side cache)."""
    s = torch.cuda.current_stream()
    fn()
    ts = []
    for _ in range(reps):
        junk.fill_(0)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        fn()
        e1.record(s)
        e1.synchronize()
        ts.append(e0.elapsed_time(e1))
    return float(np.median(ts))


def _time(fn, reps: int, warmup: int) -> float:
    for _ in range(warmup):
        fn()
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(plan, er, ur, k0, precision: str = "mixed", out=None, reps: int = 5, warmup: int = 2,
         use_cache: bool = True, candidates=CANDIDATES, windows=WINDOWS,
         rounds: int = 2, cold: bool = False, flush_bytes: int = 1 << 30,
         prefetch=PREFETCH) -> dict:
    """Pick the fastest (window, tile, prefetch) choice for
    ``plan.assemble_a(er, ur, k0, precision)``.

    Returns ``{"nb": tile, "window": window, "prefetch": chunk_tets, "ms":
    {"window/tile[/pf]": ms}, "cached": bool, "key": key}`` and leaves the plan set to the
    best choice. ``windows=None`` keeps the plan's window, ``prefetch=None`` the plan's
    prefetch setting. Every candidate is timed in
    ``rounds`` interleaved passes and keeps its fastest pass: single trials of one
    candidate vary by a few percent, as much as the candidates differ.

    ``cold=True`` times single launches that each follow a ``flush_bytes`` device fill, for
    callers that run other work between assemblies: back-to-back launches of a mesh whose
    inputs fit the 256 MB memory-side cache find them there, and the best window differs
    (a 2.8M-tet mesh: 512-tet windows warm, 1,024 cold).
    """
    er_m, ur_m = plan.material(er), plan.material(ur)
    k0a = np.atleast_1d(np.asarray(k0, np.float64))
    key = tuning_key(plan, er_m, ur_m, int(k0a.shape[0]), precision)
    if windows is None:
        windows = (plan.get_window(),)
    else:
        key += "|win"
    if prefetch is None:
        prefetch = (plan.get_prefetch(),)
    else:
        key += "|pf"
    if cold:
        key += "|cold"
    cache = _load() if use_cache else {}
    if use_cache and key in cache:
        c = cache[key]
        plan.set_window(int(c.get("window", plan.get_window())))
        plan.set_tuning(int(c["nb"]))
        plan.set_prefetch(int(c.get("prefetch", plan.get_prefetch())))
        return {"nb": int(c["nb"]), "window": plan.get_window(),
                "prefetch": plan.get_prefetch(), "ms": c.get("ms", {}), "cached": True,
                "key": key}
    if out is None:
        from .assembly import _out_dtype, _prec
        out = torch.empty((k0a.shape[0], plan.nnz_coo), dtype=_out_dtype(_prec(precision)),
                          device=plan.device)
    k0arg = float(k0a[0]) if k0a.shape[0] == 1 else k0a
    target = out[0] if (k0a.shape[0] == 1 and out.dim() == 2) else out
    junk = torch.empty(flush_bytes, dtype=torch.uint8, device=plan.device) if cold else None
    times = {}
    for _ in range(max(1, rounds)):
        for w in windows:
            plan.set_window(w)
            for pf in prefetch:
                plan.set_prefetch(pf)
                for nb in candidates:
                    plan.set_tuning(nb)
                    fn = lambda: plan.assemble_a(er_m, ur_m, k0arg, precision, out=target)  # noqa: E731
                    t = _time_cold(fn, reps, junk) if cold else _time(fn, reps, warmup)
                    times[(w, nb, pf)] = min(t, times.get((w, nb, pf), t))
    del junk
    best_w, best, best_pf = min(times, key=times.get)
    plan.set_window(best_w)
    plan.set_tuning(best)
    plan.set_prefetch(best_pf)
    ms = {f"{w}/{nb}" + (f"/pf{pf}" if pf else ""): v for (w, nb, pf), v in times.items()}
    if use_cache:
        cache = _load()
        cache[key] = {"nb": best, "window": best_w, "prefetch": best_pf, "ms": ms}
        _save(cache)
    return {"nb": best, "window": best_w, "prefetch": best_pf, "ms": ms, "cached": False,
            "key": key}
