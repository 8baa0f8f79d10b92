"""Autotuner for the assembly kernels' launch shape.

The tunable is the number of tets per workgroup tile (32, 64 or 128): it trades the
latency of the per-tile geometry phase and the tail of partly filled tiles against
occupancy, and the best value depends on the precision, the output mode (one k0 or a
sweep), the material kinds and the mesh size.
``tune`` times each candidate with HIP events on the launch stream, applies the fastest
to the plan and remembers it in a JSON cache keyed by all of the above, so later plans
of the same kind skip the timing.

Cache file: ``$EMERGE_AMD_TUNE_CACHE`` or ``~/.cache/emerge_amd/autotune.json``.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np
import torch

CANDIDATES = (32, 64, 128)


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
         use_cache: bool = True, candidates=CANDIDATES) -> dict:
    """Pick the fastest tile size for ``plan.assemble_a(er, ur, k0, precision)``.

    Returns ``{"nb": best, "ms": {nb: ms}, "cached": bool, "key": key}`` and leaves the
    plan set to the best value.
    """
    er_m, ur_m = plan.material(er), plan.material(ur)
    k0a = np.atleast_1d(np.asarray(k0, np.float64))
    key = tuning_key(plan, er_m, ur_m, int(k0a.shape[0]), precision)
    cache = _load() if use_cache else {}
    if use_cache and key in cache:
        plan.set_tuning(int(cache[key]["nb"]))
        return {"nb": int(cache[key]["nb"]), "ms": cache[key].get("ms", {}), "cached": True,
                "key": key}
    if out is None:
        dt = torch.complex128 if precision == "fp64" else torch.complex64
        out = torch.empty((k0a.shape[0], plan.nnz_coo), dtype=dt, device=plan.device)
    k0arg = float(k0a[0]) if k0a.shape[0] == 1 else k0a
    target = out[0] if (k0a.shape[0] == 1 and out.dim() == 2) else out
    times = {}
    for nb in candidates:
        plan.set_tuning(nb)
        times[nb] = _time(lambda: plan.assemble_a(er_m, ur_m, k0arg, precision, out=target),
                          reps, warmup)
    best = min(times, key=times.get)
    plan.set_tuning(best)
    if use_cache:
        cache = _load()
        cache[key] = {"nb": best, "ms": {str(k): v for k, v in times.items()}}
        _save(cache)
    return {"nb": best, "ms": times, "cached": False, "key": key}
