#!/usr/bin/env python3
"""Nedelec-2 curl-curl system assembly throughput on one MI355X.

Workload: unit-box Delaunay mesh of 400k seeded uniform points plus a jittered boundary
grid (about 2.6M tets), per-tet complex eps_r = Re(1 - j tan_d) with Re ~ U[1,10],
tan_d ~ U[0,0.02], per-tet diagonal mu_r ~ U[1,2]^3, f = 10 GHz. One step assembles the
COO values of A = K - k0^2 M for every tet in one kernel launch; inputs are resident in
HBM. The headline ``value`` is the mixed-precision kernel (FP64 geometry, FP32
contraction, complex64 output); the plain FP64 kernel (complex128 output) and the
relative Frobenius error of mixed against FP64 are reported beside it, with the CPU
restatement (oracle) timed on a bounded sample of the same mesh.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Nédélec tets assembled/s + HBM GB/s: mixed-FP32 vs FP64 at ≤1e-6 rel err"
HBM_PEAK_GBS = 8000.0
C0 = 299792458.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bytes_per_tet(out_bytes_per_entry: int, nodes_per_tet: float, mat_bytes: int = 64) -> float:
    """Algorithmic HBM bytes per tet of the fused A kernel.

    node ids 16 (4 x int32) + node coordinates 24 * Nn/T (3 x f64, each node read once)
    + materials (default: eps_r one complex128 = 16, mu_r diagonal 3 complex128 = 48)
    + 400 output entries.
    """
    return 16 + 24 * nodes_per_tet + mat_bytes + 400 * out_bytes_per_entry


def time_launches(fn, steps: int, warmup: int):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall


def cold_cache_launches(fn, reps: int = 8, flush_bytes: int = 1 << 30) -> float:
    """Median time of single launches, each after a ``flush_bytes`` device fill on the same
    stream, so the launch finds none of its inputs in the memory-side (Infinity) cache;
    back-to-back launches of one mesh whose inputs fit that cache (256 MB) find them there."""
    import numpy as np
    import torch
    junk = torch.empty(flush_bytes // 4, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    fn()
    ts = []
    for _ in range(reps):
        junk.fill_(1.0)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        ts.append(e0.elapsed_time(e1))
    del junk
    return float(np.median(ts))


def rel_err_chunked(a32, a64, chunk=1 << 26) -> float:
    import torch
    num = 0.0
    den = 0.0
    for s in range(0, a64.numel(), chunk):
        x = a64[s:s + chunk]
        d = a32[s:s + chunk].to(torch.complex128) - x
        num += float(torch.sum(d.real ** 2 + d.imag ** 2))
        den += float(torch.sum(x.real ** 2 + x.imag ** 2))
    return (num / den) ** 0.5


def tuning_record(r: dict) -> dict:
    """The autotuner's choice and the fastest time it measured per window, without and
    with prefetched launches."""
    best, best_pf = {}, {}
    for k, ms in r["ms"].items():
        w = k.split("/")[0]
        d = best_pf if "/pf" in k else best
        d[w] = min(ms, d.get(w, ms))
    return {"tile": r["nb"], "window": r["window"], "prefetch_tets": r.get("prefetch", 0),
            "best_ms_by_window": best, "best_ms_by_window_prefetched": best_pf}


def apply_tuning(plan, rec: dict) -> None:
    plan.set_window(rec["window"])
    plan.set_tuning(rec["tile"])
    plan.set_prefetch(rec["prefetch_tets"])


def host_cpu_info() -> dict:
    """The host's CPU counts: physical cores (distinct core/socket pairs of ``lscpu -p``),
    logical CPUs (``nproc``), the CPUs this process may run on, and the thread share a
    one-GPU job takes (at most 16 threads per GPU on a shared GPU host)."""
    import subprocess
    info = {"logical_cpus": os.cpu_count() or 1}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True,
                                           check=True).stdout.strip())
    except Exception:
        info["nproc"] = None
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True,
                             check=True).stdout
        info["physical_cores"] = len({ln for ln in out.splitlines()
                                      if ln and not ln.startswith("#")})
    except Exception:
        info["physical_cores"] = None
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except Exception:
        info["affinity_cpus"] = info["logical_cpus"]
    info["job_threads"] = max(1, min(16, info["affinity_cpus"], info["physical_cores"] or 16))
    return info


def cpu_baseline(field, er_ref, ur_ref, seconds: float, chunk: int, nthreads: int = 0):
    """The oracle (C/OpenMP restatement of _matrix_builder, K and M plus int64 rows/cols
    like the reference) on the first ``chunk`` tets of the mesh, repeated for about
    ``seconds``; returns tets/s and the thread count (default: the job's thread share,
    ``host_cpu_info()["job_threads"]``)."""
    import numpy as np
    import oracle
    nthreads = nthreads or host_cpu_info()["job_threads"]
    m = field.mesh
    ids = np.arange(min(chunk, field.n_tets), dtype=np.int64)
    args = (m.nodes, m.tets, m.tris, m.edges, field.tet_to_field, m.tet_to_edge, m.tet_to_tri,
            ur_ref, er_ref, ids, field.dofcodes3d)
    oracle.matrix_builder(*args, nthreads=nthreads)  # warm
    done, t0 = 0, time.perf_counter()
    while True:
        oracle.matrix_builder(*args, nthreads=nthreads)
        done += ids.shape[0]
        el = time.perf_counter() - t0
        if el >= seconds and done >= 3 * ids.shape[0]:
            break
    return done / el, nthreads, done, el


def csc_per_frequency(plan, eps, mu, k0, steps: int, warmup: int) -> dict:
    """The per-frequency CSC matrix of the sweep (assembler.py:494, once per frequency in
    the reference): with K and M cached, their CSC sums are formed once and each frequency
    is one streaming pass A = K_csc - k0^2 M_csc (em_csc_freq_combine: 2 c bytes read and c
    written per CSC entry). Beside it, the fused column gather over the COO values that the
    reference's order implies, to_csc(E - k0^2 B). complex64 (mixed) and complex128 (the
    reference's dtype); scalar eps_r and diagonal mu_r give symmetric blocks (mirrored
    gather)."""
    import torch
    from emerge_amd.assembly import CscMap, FreqCsc, csc_freq_combine
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cm = CscMap(plan)
    torch.cuda.synchronize()
    res = {"nnz": cm.nnz, "map_ms_one_time": (time.perf_counter() - t0) * 1e3,
           "kernel": "csc_freq_kernel (em_csc_freq_combine)"}
    for prec, c in (("mixed", 8), ("fp64", 16)):
        K, M = plan.assemble_km(eps, mu, prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fc = FreqCsc(cm, K, M, block_symmetric=True)
        torch.cuda.synchronize()
        sums_ms = (time.perf_counter() - t0) * 1e3
        out = torch.empty(cm.nnz, dtype=fc.K.dtype, device="cuda")
        ms, _ = time_launches(lambda: csc_freq_combine(fc.K, fc.M, k0, out=out), steps, warmup)
        ref = cm.to_csc_combine(K, M, k0, block_symmetric=True)
        torch.cuda.synchronize()
        diff = float(torch.linalg.norm((out - ref).to(torch.complex128)) /
                     torch.linalg.norm(ref.to(torch.complex128)))
        del ref
        ms_gather, _ = time_launches(
            lambda: cm.to_csc_combine(K, M, k0, out=out, block_symmetric=True),
            max(2, steps // 2), 1)
        # A straight into CSC, without the COO arrays (em_ned2_assemble_csc), beside the
        # two-step assemble_a + to_csc it equals bit for bit
        ms_direct, _ = time_launches(
            lambda: plan.assemble_csc(cm, eps, mu, k0, "A", prec, out=out),
            max(2, steps // 2), 1)
        a_coo = plan.assemble_a(eps, mu, k0, prec)
        ms_two_step, _ = time_launches(
            lambda: cm.to_csc(plan.assemble_a(eps, mu, k0, prec, out=a_coo), out=out,
                              block_symmetric=True), max(2, steps // 2), 1)
        direct_equal = bool(torch.equal(plan.assemble_csc(cm, eps, mu, k0, "A", prec),
                                        cm.to_csc(a_coo, block_symmetric=True)))
        del a_coo
        b = 3 * cm.nnz * c
        res[prec] = {"ms_per_frequency": ms, "bytes_per_frequency": b,
                     "gb_per_s": b / (ms * 1e-3) / 1e9,
                     "roofline_frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "ms_one_time_csc_sums_of_K_and_M": sums_ms,
                     "ms_per_frequency_fused_coo_gather": ms_gather,
                     "rel_diff_vs_fused_gather": diff,
                     "ms_direct_csc_A_no_coo": ms_direct,
                     "ms_assemble_a_then_to_csc": ms_two_step,
                     "direct_equals_two_step": direct_equal}
        log(f"csc per frequency ({prec}): {ms:.3f} ms combine, {ms_gather:.3f} ms fused gather")
        del K, M, fc, out
        torch.cuda.empty_cache()
    cm.close()
    return res


def _mats(T, rng, mu_kind="diag"):
    import numpy as np
    import torch
    eps = torch.from_numpy(rng.uniform(1, 10, T) * (1 - 1j * rng.uniform(0, 0.02, T)))
    if mu_kind == "diag":
        mu = torch.from_numpy(rng.uniform(1, 2, (T, 3)).astype(np.complex128))
    else:
        mu = torch.ones(T, dtype=torch.complex128)
    return eps, mu


def suite_sweep(args) -> dict:
    """64 frequencies over 8-12 GHz on a ~500k-tet Delaunay mesh, eps_r = mu_r = 1, all
    A_f in one graph launch (complex64 COO, 102 GB)."""
    import numpy as np
    import torch
    from emerge_amd.assembly import Ned2Plan
    from emerge_amd.graph import SweepGraph
    from emerge_amd.mesh import Nedelec2Field, delaunay_mesh_with_tets
    f = Nedelec2Field(delaunay_mesh_with_tets(500_000, seed=args.seed + 3))
    plan = Ned2Plan.from_field(f)
    T = plan.n_tets
    k0s = 2 * np.pi * np.linspace(8e9, 12e9, 64) / C0
    one = plan.material(1.0)
    sg = SweepGraph(plan, one, one, k0s, "mixed")
    log(f"sweep: {T} tets x 64 frequencies, output {sg.out.numel() * 8 / 1e9:.1f} GB")
    ms_graph, _ = time_launches(lambda: sg.replay(), 3, 1)
    ms_eager, _ = time_launches(lambda: plan.assemble_a(one, one, k0s, "mixed", out=sg.out), 3, 1)
    ref = plan.assemble_a(one, one, float(k0s[17]), "mixed")
    sg.replay()
    torch.cuda.synchronize()
    diff = float(torch.linalg.norm(ref - sg.out[17]) / torch.linalg.norm(ref))
    del sg, ref
    torch.cuda.empty_cache()
    out_bytes = 64 * T * 400 * 8
    res = {"n_tets": T, "n_freq": 64, "output": "complex64 COO per frequency",
           "ms_per_sweep_graph": ms_graph, "ms_per_sweep_eager": ms_eager,
           "tet_freqs_per_s": 64 * T / (ms_graph * 1e-3),
           "write_gb_per_s": out_bytes / (ms_graph * 1e-3) / 1e9,
           "rel_diff_replay_vs_single_launch": diff}
    # the reference's product: one CSC matrix per frequency (assembler.py:494), K and M shared
    from emerge_amd.assembly import CscMap, FreqCsc, csc_freq_combine
    from emerge_amd.graph import CscSweepGraph
    cm = CscMap(plan)
    cg = CscSweepGraph(plan, cm, one, one, k0s, "mixed")
    log(f"csc sweep: nnz {cm.nnz}, output {cg.out.numel() * 8 / 1e9:.1f} GB")
    ms_csc_graph, _ = time_launches(lambda: cg.replay(), 3, 1)
    fc = FreqCsc(cm, cg.K, cg.M, block_symmetric=True)
    ms_combine, _ = time_launches(lambda: csc_freq_combine(fc.K, fc.M, k0s, out=cg.out), 3, 1)
    ref = cm.to_csc_combine(cg.K, cg.M, float(k0s[17]), block_symmetric=True)
    cg.replay()
    torch.cuda.synchronize()
    dcsc = float(torch.linalg.norm((cg.out[17] - ref).to(torch.complex128)) /
                 torch.linalg.norm(ref.to(torch.complex128)))
    # algorithmic bytes of the whole CSC sweep: K and M COO blocks written (2 x 400 x 8 per
    # tet) and read by the two column gathers (+ 2 B row position per contributor), the CSC
    # sums written and read once (2 x 2 x nnz x 8), 64 A_f written (64 x nnz x 8)
    b_sweep = 2 * plan.nnz_coo * 8 + 2 * plan.nnz_coo * 10 + 4 * cm.nnz * 8 + 64 * cm.nnz * 8
    b_combine = (2 + 64) * cm.nnz * 8
    res["csc"] = {"nnz": cm.nnz, "output": "complex64 CSC values per frequency",
                  "ms_per_sweep_graph": ms_csc_graph,
                  "ms_combine_64_frequencies": ms_combine,
                  "combine_gb_per_s": b_combine / (ms_combine * 1e-3) / 1e9,
                  "combine_roofline_frac": b_combine / (ms_combine * 1e-3) / 1e9 / HBM_PEAK_GBS,
                  "sweep_gb_per_s": b_sweep / (ms_csc_graph * 1e-3) / 1e9,
                  "tet_freqs_per_s": 64 * T / (ms_csc_graph * 1e-3),
                  "rel_diff_vs_fused_gather_f17": dcsc}
    del cg, fc, ref, cm
    torch.cuda.empty_cache()
    return res


def suite_ref(args) -> dict:
    """The reference-arithmetic FP64 mode (EM_PREC_FP64_REF: 11-point rule in absolute
    coordinates, curlcurl.py:285-380) against the exact-integral FP64 kernel on a ~500k-tet
    Delaunay mesh: K and M (the reference's dataE, dataB), complex128, per-tet complex
    eps_r, diagonal mu_r; and their relative Frobenius distance on A at 10 GHz."""
    import numpy as np
    import torch
    from emerge_amd.assembly import Ned2Plan
    from emerge_amd.mesh import Nedelec2Field, delaunay_mesh_with_tets
    f = Nedelec2Field(delaunay_mesh_with_tets(500_000, seed=args.seed + 3))
    plan = Ned2Plan.from_field(f)
    T = plan.n_tets
    eps, mu = _mats(T, np.random.default_rng(args.seed + 21))
    eps, mu = plan.material(eps), plan.material(mu)
    ms = {}
    for prec in ("fp64_ref", "fp64"):
        ms[prec], _ = time_launches(lambda: plan.assemble_km(eps, mu, prec), 3, 1)
        log(f"ref suite: {prec} K and M {ms[prec]:.3f} ms")
    k0 = 2 * np.pi * 10e9 / C0
    a_ref = plan.assemble_a(eps, mu, k0, "fp64_ref")
    a_ex = plan.assemble_a(eps, mu, k0, "fp64")
    gap = float(torch.linalg.norm(a_ref - a_ex) / torch.linalg.norm(a_ex))
    del a_ref, a_ex
    plan.close()
    torch.cuda.empty_cache()
    return {"n_tets": T, "output": "K and M, complex128 COO",
            "ms_fp64_ref": ms["fp64_ref"], "ms_fp64": ms["fp64"],
            "tets_per_s_fp64_ref": T / (ms["fp64_ref"] * 1e-3),
            "tets_per_s_fp64": T / (ms["fp64"] * 1e-3),
            "rel_frobenius_ref_vs_exact_A": gap}


def suite_dropin(args) -> dict:
    """The Python drop-in tet_mass_stiffness_matrices (curlcurl.py:222-270) end to end on a
    ~500k-tet Delaunay mesh: host (3,3,T) complex128 er/ur in, host complex128 dataE/dataB
    out (the reference's interface), so both PCIe directions are inside the timed call;
    beside it the device-resident assembly of the same matrices and the bare copies."""
    import numpy as np
    import torch
    from emerge_amd.curlcurl import tet_mass_stiffness_matrices
    from emerge_amd.mesh import Nedelec2Field, delaunay_mesh_with_tets
    f = Nedelec2Field(delaunay_mesh_with_tets(500_000, seed=args.seed + 3))
    T = f.n_tets
    rng = np.random.default_rng(args.seed + 23)
    er = np.zeros((3, 3, T), np.complex128)
    ur = np.zeros((3, 3, T), np.complex128)
    eps = rng.uniform(1, 10, T) * (1 - 1j * rng.uniform(0, 0.02, T))
    for d in range(3):
        er[d, d] = eps
        ur[d, d] = rng.uniform(1, 2, T)
    cond = np.array([], np.int64)
    _, _, cm = tet_mass_stiffness_matrices(f, er, ur, cond)  # plan, map and first copies
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        dE, dB, _ = tet_mass_stiffness_matrices(f, er, ur, cond, cscmap=cm)
    call_s = (time.perf_counter() - t0) / reps
    from emerge_amd.curlcurl import plan_for
    plan = plan_for(f, cond)
    er_d, ur_d = plan.material(er, layout="ref"), plan.material(ur, layout="ref")
    dev_ms, _ = time_launches(lambda: plan.assemble_km(er_d, ur_d, "fp64"), 5, 1)
    K, M = plan.assemble_km(er_d, ur_d, "fp64")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        K.cpu(), M.cpu()
    pageable_s = (time.perf_counter() - t0) / reps
    from emerge_amd.assembly import to_host
    t0 = time.perf_counter()
    for _ in range(reps):
        hK, hM = to_host(K), to_host(M)
    d2h_s = (time.perf_counter() - t0) / reps
    del hK, hM
    # the reference's next step, cscmap.to_csc(Evec - Bvec * K0**2), on host COO values
    k0 = 2 * np.pi * 10e9 / C0
    coo = dE - dB * k0 ** 2
    t0 = time.perf_counter()
    for _ in range(reps):
        A = cm.to_csc(coo)
    to_csc_s = (time.perf_counter() - t0) / reps
    nnz = A.nnz
    del A
    host = torch.from_numpy(coo)
    t0 = time.perf_counter()
    for _ in range(reps):
        d = host.to("cuda")
        torch.cuda.synchronize()
    h2d_pageable_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        d = host.pin_memory().to("cuda", non_blocking=True)
        torch.cuda.synchronize()
    h2d_pinned_s = (time.perf_counter() - t0) / reps
    del d, host, coo
    out_bytes = 2 * K.numel() * 16
    del K, M, dE, dB
    torch.cuda.empty_cache()
    return {"n_tets": T, "output": "dataE, dataB host complex128 (2 x 6,400 B/tet)",
            "s_per_call": call_s, "tets_per_s_call": T / call_s,
            "ms_device_assembly_km": dev_ms, "tets_per_s_device": T / (dev_ms * 1e-3),
            "s_d2h_copies": d2h_s, "d2h_gb_per_s": out_bytes / d2h_s / 1e9,
            "s_d2h_pageable": pageable_s, "d2h_pageable_gb_per_s": out_bytes / pageable_s / 1e9,
            "s_to_csc_call": to_csc_s, "csc_nnz": nnz,
            "s_h2d_coo_pageable": h2d_pageable_s, "s_h2d_coo_pinned_staged": h2d_pinned_s,
            "note": "fresh host arrays per call, as the reference returns; the drop-in "
                    "copies into pinned memory (s_d2h_copies), torch's .cpu() into "
                    "pageable memory (s_d2h_pageable)"}


def suite_batch(args) -> dict:
    """32 Delaunay meshes, tet counts log-uniform in [5k, 400k], per-tet complex eps_r,
    mu_r = 1, mixed precision, one launch; per-mesh error against the FP64 kernel (which
    matches the oracle to 1e-12 in the tests)."""
    import numpy as np
    import torch
    from emerge_amd import autotune
    from emerge_amd.assembly import Ned2Plan
    from emerge_amd.mesh import Nedelec2Field, delaunay_mesh_with_tets
    rng = np.random.default_rng(args.seed + 4)
    sizes = np.exp(rng.uniform(np.log(5e3), np.log(4e5), 32)).astype(int)
    fields = [Nedelec2Field(delaunay_mesh_with_tets(int(s), seed=args.seed + 100 + i))
              for i, s in enumerate(sizes)]
    plan = Ned2Plan.from_fields(fields)
    T = plan.n_tets
    eps, mu = _mats(T, rng, mu_kind="one")
    er, ur = plan.material(eps), plan.material(mu)
    k0 = 2 * np.pi * 10e9 / C0
    log(f"batch: 32 meshes, {T} tets ({sizes.min()}..{sizes.max()} requested)")
    a32 = torch.empty(plan.nnz_coo, dtype=torch.complex64, device="cuda")
    tune = autotune.tune(plan, er, ur, k0, "mixed", out=a32[None], reps=3)
    ms, _ = time_launches(lambda: plan.assemble_a(er, ur, k0, "mixed", out=a32), 10, 2)
    a64 = plan.assemble_a(er, ur, k0, "fp64")
    plan.assemble_a(er, ur, k0, "mixed", out=a32)
    torch.cuda.synchronize()
    errs = []
    for i in range(32):
        s0, s1 = 400 * int(plan.tet_offsets[i]), 400 * int(plan.tet_offsets[i + 1])
        errs.append(rel_err_chunked(a32[s0:s1], a64[s0:s1]))
    del a64, a32
    torch.cuda.empty_cache()
    b = bytes_per_tet(8, sum(f.mesh.n_nodes for f in fields) / T, mat_bytes=32)
    return {"n_meshes": 32, "n_tets": T, "tets_per_mesh": [f.n_tets for f in fields],
            "ms_per_launch": ms, "tets_per_s": T / (ms * 1e-3),
            "gb_per_s": T * b / (ms * 1e-3) / 1e9, "tuned_nb": tune["nb"], "tuned_window": tune["window"],

            "rel_err_vs_fp64_per_mesh": errs, "max_rel_err": max(errs)}


def suite_backward(args) -> dict:
    """dL/d eps_r and dL/d mu_r^-1 on a ~1M-tet mesh for L = Re(lam^H A x)."""
    import numpy as np
    import torch
    from emerge_amd.assembly import Ned2Plan
    from emerge_amd.mesh import Nedelec2Field, delaunay_mesh_with_tets
    f = Nedelec2Field(delaunay_mesh_with_tets(1_000_000, seed=args.seed + 5))
    plan = Ned2Plan.from_field(f)
    T, N = plan.n_tets, f.n_field
    rng = np.random.default_rng(args.seed + 5)
    lam = torch.from_numpy(rng.normal(size=N) + 1j * rng.normal(size=N)).cuda()
    x = torch.from_numpy(rng.normal(size=N) + 1j * rng.normal(size=N)).cuda()
    rows, cols = plan.rowcol()
    g = lam[rows] * torch.conj(x[cols])
    del rows, cols
    ge = torch.zeros(T, dtype=torch.complex128, device="cuda")
    gm = torch.zeros((T, 3), dtype=torch.complex128, device="cuda")
    from emerge_amd import _lib
    k0 = 2 * np.pi * 10e9 / C0
    fn = lambda: plan.backward(g, k0, _lib.EM_MAT_SCALAR, ge, _lib.EM_MAT_DIAG, gm)  # noqa: E731
    ms, _ = time_launches(fn, 10, 2)
    log(f"backward: {T} tets, {ms:.3f} ms")
    # full (T, 3, 3) leaves (antisymmetric parts included)
    gef = torch.zeros((T, 9), dtype=torch.complex128, device="cuda")
    gmf = torch.zeros((T, 9), dtype=torch.complex128, device="cuda")
    fnf = lambda: plan.backward(g, k0, _lib.EM_MAT_FULL, gef, _lib.EM_MAT_FULL, gmf)  # noqa: E731
    ms_full, _ = time_launches(fnf, 10, 2)
    log(f"backward full tensors: {ms_full:.3f} ms")
    bf = 400 * 16 + 16 + 24 * f.mesh.n_nodes / T + 2 * 144
    b = 400 * 16 + 16 + 24 * f.mesh.n_nodes / T + 16 + 48
    return {"n_tets": T, "ms_per_launch": ms, "tets_per_s": T / (ms * 1e-3),
            "gb_per_s": T * b / (ms * 1e-3) / 1e9, "bytes_per_tet": b,
            "leaves": "eps_r (T,) complex, mu_r^-1 (T,3) complex",
            "full_tensor_leaves": {"ms_per_launch": ms_full, "tets_per_s": T / (ms_full * 1e-3),
                                   "gb_per_s": T * bf / (ms_full * 1e-3) / 1e9,
                                   "bytes_per_tet": bf}}


def suite_csc_pec(args) -> dict:
    """COO -> CSC and PEC elimination on a ~500k-tet Delaunay mesh (all outer faces PEC):
    one-time CSC map / PEC mask / restricted map builds, then per frequency the fused
    to_csc(E - k0^2 B) of the full matrix, the free-DOF block selected from it, and the
    restricted map that writes the block straight from the COO values (complex64).
    The CPU line is SciPy's A[:, solve_ids][solve_ids, :] (the reference's own call,
    solver.py:2071) on the same matrix."""
    import numpy as np
    import torch
    from emerge_amd.assembly import CscMap, Ned2Plan
    from emerge_amd.mesh import Nedelec2Field, delaunay_mesh_with_tets
    from emerge_amd.pec import pec_solve_ids, select_csc, to_scipy
    f = Nedelec2Field(delaunay_mesh_with_tets(500_000, seed=args.seed + 7))
    plan = Ned2Plan.from_field(f)
    T, N = plan.n_tets, f.n_field
    eps, mu = _mats(T, np.random.default_rng(args.seed + 7))
    K, M = plan.assemble_km(eps, mu, "mixed")
    k0 = 2 * np.pi * 10e9 / C0

    def once(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, (time.perf_counter() - t0) * 1e3

    cm, ms_map = once(lambda: CscMap(plan))
    tris = f.mesh.boundary_tris()
    sel, ms_mask = once(lambda: pec_solve_ids(plan, f, [], tris))
    sub, ms_restrict = once(lambda: cm.restrict(sel))
    log(f"csc/pec: {T} tets, N={N}, nnz={cm.nnz}, free DOFs {sel.n_solve}, block nnz {sub.nnz}")
    full = torch.empty(cm.nnz, dtype=torch.complex64, device="cuda")
    blk = torch.empty(sub.nnz, dtype=torch.complex64, device="cuda")
    ex = cm.export("indptr", "indices")
    ms_full, _ = time_launches(lambda: cm.to_csc_combine(K, M, k0, out=full), 10, 2)
    ms_sel, _ = time_launches(lambda: select_csc(ex["indptr"], ex["indices"], full, sel), 5, 1)
    ms_sub, _ = time_launches(lambda: sub.to_csc_combine(K, M, k0, out=blk), 10, 2)
    # scalar eps_r and diagonal mu_r: symmetric blocks, mirrored-read gather
    ms_full_m, _ = time_launches(
        lambda: cm.to_csc_combine(K, M, k0, out=full, block_symmetric=True), 10, 2)
    ms_sub_m, _ = time_launches(
        lambda: sub.to_csc_combine(K, M, k0, out=blk, block_symmetric=True), 10, 2)
    same_m = bool(torch.equal(sub.to_csc_combine(K, M, k0),
                              sub.to_csc_combine(K, M, k0, block_symmetric=True)))
    # the full map's default is the per-column incidence gather; the per-entry segment
    # gather (what restricted maps use) beside it
    ms_full_seg, _ = time_launches(
        lambda: cm.to_csc_combine(K, M, k0, out=full, segment_gather=True), 10, 2)
    ms_full_seg_m, _ = time_launches(
        lambda: cm.to_csc_combine(K, M, k0, out=full, block_symmetric=True,
                                  segment_gather=True), 10, 2)
    same_inc = bool(torch.equal(
        cm.to_csc_combine(K, M, k0, block_symmetric=True),
        cm.to_csc_combine(K, M, k0, block_symmetric=True, segment_gather=True)))
    # blocks that are not symmetric (full anisotropic tensors) cannot use the mirrored
    # reads; the staged gather writes the contributions block-transposed first
    work = torch.empty(plan.nnz_coo, dtype=torch.complex128, device="cuda")
    ms_staged, _ = time_launches(lambda: cm.to_csc_staged(K, M, k0, work=work, out=full), 10, 2)
    same_staged = bool(torch.equal(cm.to_csc_staged(K, M, k0, work=work),
                                   cm.to_csc_combine(K, M, k0)))
    del work
    # complex128 (the reference's dtype): the same gathers on FP64 K and M
    K64, M64 = plan.assemble_km(eps, mu, "fp64")
    full64 = torch.empty(cm.nnz, dtype=torch.complex128, device="cuda")
    ms64_m, _ = time_launches(
        lambda: cm.to_csc_combine(K64, M64, k0, out=full64, block_symmetric=True), 10, 2)
    ms64, _ = time_launches(lambda: cm.to_csc_combine(K64, M64, k0, out=full64), 10, 2)
    del K64, M64, full64
    ref_blk = select_csc(ex["indptr"], ex["indices"], cm.to_csc_combine(K, M, k0), sel)[2]
    sub.to_csc_combine(K, M, k0, out=blk)
    torch.cuda.synchronize()
    same = bool(torch.equal(ref_blk, blk))
    del ref_blk
    # algorithmic bytes of a gather: per COO contributor perm (8) + E, B (8 + 8), per CSC
    # entry seg (8) + value out (8). Contributors of the block: COO entries whose CSC entry
    # has a free row and a free column.
    tgt = cm.export("target_csc")["target_csc"]
    keep = sel.keep.bool()
    col = torch.repeat_interleave(torch.arange(N, device="cuda"), torch.diff(ex["indptr"]))
    kept_q = keep[ex["indices"]] & keep[col]
    n_contrib = int(kept_q[tgt].sum())
    del tgt, col, kept_q
    # the oracle's restatement of CSCMapping (pattern, then the per-frequency scatter) on
    # the first 50k tets' COO entries, one host thread, scaled to the whole mesh
    import oracle
    n_cpu = min(T, 50_000)
    rows_all, cols_all = plan.rowcol()
    r_h = rows_all[:400 * n_cpu].cpu().numpy()
    c_h = cols_all[:400 * n_cpu].cpu().numpy()
    del rows_all, cols_all
    d_h = (K[:400 * n_cpu] - (k0 ** 2) * M[:400 * n_cpu]).to(torch.complex128).cpu().numpy()
    t0 = time.perf_counter()
    ocm = oracle.CSCMap(r_h, c_h, N)
    cpu_map_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    ocm.to_csc_data(d_h)
    cpu_csc_ms = (time.perf_counter() - t0) * 1e3
    del ocm, r_h, c_h, d_h
    A_h = to_scipy(ex["indptr"], ex["indices"], full.to(torch.complex128), N)
    ids = sel.solve_ids.cpu().numpy()
    t0 = time.perf_counter()
    blk_h = A_h[:, ids][ids, :]
    cpu_ms = (time.perf_counter() - t0) * 1e3
    assert blk_h.nnz == sub.nnz
    del A_h, blk_h
    res = {"n_tets": T, "n_dofs": N, "nnz": cm.nnz, "n_solve": sel.n_solve,
           "block_nnz": sub.nnz,
           "one_time_ms": {"csc_map": ms_map, "pec_mask_and_solve_ids": ms_mask,
                           "restricted_map": ms_restrict},
           "per_frequency_ms": {"to_csc_combine_full": ms_full, "select_block_from_full": ms_sel,
                                "restricted_to_csc_combine": ms_sub,
                                "to_csc_combine_full_mirrored": ms_full_m,
                                "restricted_to_csc_combine_mirrored": ms_sub_m,
                                "to_csc_combine_full_segment_gather": ms_full_seg,
                                "to_csc_combine_full_segment_gather_mirrored": ms_full_seg_m,
                                "to_csc_combine_full_staged": ms_staged},
           "staged_equals_plain": same_staged,
           "incidence_equals_segment_gather": same_inc,
           "restricted_equals_selected_full": same,
           "mirrored_equals_plain": same_m,
           "cpu_scipy_select_ms": cpu_ms,
           "cpu_oracle_csc": {"sample_tets": n_cpu, "threads": 1,
                              "map_ms": cpu_map_ms, "to_csc_ms": cpu_csc_ms,
                              "map_ms_scaled": cpu_map_ms * T / n_cpu,
                              "to_csc_ms_scaled": cpu_csc_ms * T / n_cpu},
           "block_coo_contributors": n_contrib,
           "fp64_per_frequency_ms": {"to_csc_combine_full_mirrored": ms64_m,
                                     "to_csc_combine_full": ms64},
           # algorithmic bytes of the incidence gather: per COO contributor its row
           # position (2) + E and B (8 + 8 complex64, 16 + 16 complex128), per CSC entry
           # the value out (8 / 16)
           "full_mirrored_gb_per_s":
               (plan.nnz_coo * 18 + cm.nnz * 8) / (ms_full_m * 1e-3) / 1e9,
           "full_mirrored_fp64_gb_per_s":
               (plan.nnz_coo * 34 + cm.nnz * 16) / (ms64_m * 1e-3) / 1e9,
           "restricted_mirrored_gb_per_s":
               (n_contrib * 18 + sub.nnz * 8) / (ms_sub_m * 1e-3) / 1e9,
           # the segment gather's: perm 8 + E, B 8 + 8 per contributor, seg 8 + out 8 per
           # entry
           "full_segment_gather_gb_per_s":
               (plan.nnz_coo * 24 + cm.nnz * 16) / (ms_full_seg * 1e-3) / 1e9}
    del full, blk, K, M, cm, sub
    torch.cuda.empty_cache()
    return res


def suite_surface_interp(args) -> dict:
    """Robin/port surface assembly and field interpolation on a ~500k-tet Delaunay mesh:
    the 8x8 blocks of every outer face (absorbing boundary), their CSC positions, the
    per-frequency add into the CSC values, a port excitation vector; point location and
    interpolation of a random solution at 2M random points. CPU lines: the oracle's
    restatements of compute_bc_entries and ned2_tet_interp (one thread) on samples."""
    import numpy as np
    import torch
    import oracle
    from oracle import interp as oi
    from oracle import robin as orob
    from emerge_amd.assembly import CscMap, Ned2Plan
    from emerge_amd.interp import interp, locate
    from emerge_amd.mesh import Nedelec2Field, delaunay_mesh_with_tets
    from emerge_amd.robin import RobinSurface
    f = Nedelec2Field(delaunay_mesh_with_tets(500_000, seed=args.seed + 8))
    plan = Ned2Plan.from_field(f)
    T = plan.n_tets
    m = f.mesh
    surf = m.boundary_tris()
    rs = RobinSurface(plan, f, surf)
    unit = rs.blocks(1.0)
    ms_blocks, _ = time_launches(lambda: rs.blocks(1.0, out=unit), 10, 2)
    cm = CscMap(plan)
    data = torch.zeros(cm.nnz, dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pos = rs.positions(cm)
    torch.cuda.synchronize()
    ms_pos = (time.perf_counter() - t0) * 1e3
    ms_add, _ = time_launches(lambda: rs.add_to_csc(data, pos, unit, 1j * 150.0), 10, 2)
    U = torch.randn((3, 6, len(surf)), dtype=torch.complex128, device="cuda")
    bvec = torch.zeros(f.n_field, dtype=torch.complex128, device="cuda")
    ms_force, _ = time_launches(lambda: rs.force(U, bvec=bvec), 5, 1)
    ns = min(len(surf), 20000)
    t0 = time.perf_counter()
    orob.bc_entries(m.nodes, m.tris, surf[:ns], 1.0)
    cpu_blocks = (time.perf_counter() - t0) / ns * len(surf) * 1e3
    # interpolation
    rng = np.random.default_rng(args.seed + 8)
    npts = 2_000_000
    pts = torch.as_tensor(rng.uniform(0, 1, (3, npts)), device="cuda")
    sol = torch.as_tensor(rng.normal(size=f.n_field) + 1j * rng.normal(size=f.n_field),
                          device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tet = locate(plan, pts)
    torch.cuda.synchronize()
    ms_locate = (time.perf_counter() - t0) * 1e3
    found = int((tet >= 0).sum())
    ms_interp, _ = time_launches(lambda: interp(plan, pts, sol, tet), 10, 2)
    ms_curl, _ = time_launches(lambda: interp(plan, pts, sol, tet, curl=True), 10, 2)
    ncpu = 100_000
    tet_h = tet[:ncpu].cpu().numpy()
    t0 = time.perf_counter()
    oi.interp(f, pts[:, :ncpu].cpu().numpy(), sol.cpu().numpy(), tet_h)
    cpu_interp = (time.perf_counter() - t0) / ncpu * npts * 1e3
    del data, pos, unit, cm, pts, sol, tet
    torch.cuda.empty_cache()
    return {"n_tets": T, "n_faces": len(surf), "n_points": npts, "points_located": found,
            "ms": {"robin_blocks": ms_blocks, "csc_positions_one_time": ms_pos,
                   "robin_add_per_frequency": ms_add, "port_force": ms_force,
                   "locate": ms_locate, "interp_E": ms_interp, "interp_curl": ms_curl},
            "faces_per_s": len(surf) / (ms_blocks * 1e-3),
            "points_per_s_interp": npts / (ms_interp * 1e-3),
            "cpu_ms_oracle_blocks_1thread_scaled": cpu_blocks,
            "cpu_ms_oracle_interp_scaled": cpu_interp,
            "note": "CPU figures: the oracle on 20k faces / 100k points, scaled to the full "
                    "counts; interp oracle uses OpenMP threads"}


def suite_materials(args) -> dict:
    """Material preparation of assemble_freq_matrix (assembler.py:439-469): er(1 - j tand)
    - j cond/(w eps0), the conductor ids and the frequency-dependence flag, on (3,3,T)
    complex128 tensors with T = the bench mesh's tet count and 1% copper tets. One call =
    the kernel, the id compaction and the host copy of the id count (the call blocks); the
    bandwidth counts 576 B per tet (3 x 144 B read, 144 B written). CPU line: the oracle
    (the NumPy expression plus the reference's per-tet loop) on 100k tets, scaled."""
    import numpy as np
    import torch
    from oracle import materials as om
    from emerge_amd.materials import prepare_materials
    T = 2_815_239
    g = torch.Generator(device="cuda").manual_seed(args.seed + 11)

    def rand(scale, shift=0.0):
        re = torch.rand((3, 3, T), generator=g, device="cuda", dtype=torch.float64)
        im = torch.rand((3, 3, T), generator=g, device="cuda", dtype=torch.float64)
        return torch.complex(re * scale + shift, im * scale * 1e-3)

    er, tand = rand(9.0, 1.0), rand(0.02)
    cond = torch.zeros((3, 3, T), dtype=torch.complex128, device="cuda")
    cu = torch.arange(0, T, 100, device="cuda")
    for d in range(3):
        cond[d, d, cu] = 5.8e7
    out = torch.empty_like(er)
    f = 10e9
    _, ids, fd = prepare_materials(er, tand, cond, f, out=out)
    ms, _ = time_launches(lambda: prepare_materials(er, tand, cond, f, out=out), 10, 2)
    n = 100_000
    e_np, t_np, c_np = (x[:, :, :n].cpu().numpy() for x in (er, tand, cond))
    t0 = time.perf_counter()
    _, ids_cpu, _ = om.prepare_materials(e_np, t_np, c_np, f)
    cpu_ms = (time.perf_counter() - t0) * 1e3 * T / n
    return {"tets": T, "conductor_tets": int(ids.size), "frequency_dependent": fd,
            "ms_per_call": ms, "gb_per_s_call": T * 576 / (ms * 1e-3) / 1e9,
            "ids_equal_oracle_on_sample": bool(np.array_equal(ids[ids < n], ids_cpu)),
            "cpu_ms_oracle_1thread_scaled": cpu_ms,
            "note": "call time includes the id compaction's scan, its temporaries and the "
                    "blocking copy of the count"}


def suite_volume_step(args) -> dict:
    """The whole volume step of assemble_freq_matrix (assembler.py:423-494) per frequency
    on the bench mesh, through emerge_amd.assembler.VolumeAssembler with (3,3,T) complex128
    device inputs (lossy diagonal er, diagonal ur, tand; no conductor tets): material
    preparation, element matrices, CSC value array of A. Two material sets: conductivity 0
    (nothing frequency dependent: K, M and their CSC sums cached after the first frequency,
    then one combine per frequency) and a finite conductivity in 1% of the tets
    (frequency dependent: the first frequency forms A in the element kernel and gathers
    it once, beside it the K/M assembly plus fused combine that replaced it; the second
    finds the same arrays and assembles the affine decomposition, K, M(er (1 - j tand)) and
    M(cond) in CSC form; the rest are one three-array combine each). Wall time per call,
    synchronised, median of 5 frequencies after a first one."""
    import numpy as np
    import torch
    from emerge_amd.assembler import VolumeAssembler
    from emerge_amd.mesh import Nedelec2Field, delaunay_box_mesh
    f = Nedelec2Field(delaunay_box_mesh(args.points, seed=args.seed))
    T = f.n_tets
    g = torch.Generator(device="cuda").manual_seed(args.seed + 31)
    er = torch.zeros((3, 3, T), dtype=torch.complex128, device="cuda")
    ur = torch.zeros_like(er)
    tand = torch.zeros_like(er)
    for d in range(3):
        er[d, d] = torch.rand(T, generator=g, device="cuda", dtype=torch.float64) * 9 + 1
        ur[d, d] = torch.rand(T, generator=g, device="cuda", dtype=torch.float64) + 1
        tand[d, d] = torch.rand(T, generator=g, device="cuda", dtype=torch.float64) * 0.02
    cond0 = torch.zeros_like(er)
    cond1 = torch.zeros_like(er)
    lossy = torch.arange(0, T, 100, device="cuda")
    for d in range(3):
        cond1[d, d, lossy] = 1e3
    freqs = np.linspace(8e9, 12e9, 6)
    res = {"n_tets": T}
    for prec in ("fp64", "mixed"):
        for name, cond in (("frequency_independent", cond0), ("frequency_dependent", cond1)):
            km = {}
            if name == "frequency_dependent":
                # the path the first frequency's replaced: both COO arrays, then the fused
                # combine (timed alone, before any cache holds memory)
                va = VolumeAssembler(f, prec)
                plan, er_m, ur_m, _, _ = va._prepare(er, ur, tand, cond, float(freqs[-1]))
                cm = va._cscmap(plan)
                k0 = 2 * np.pi * freqs[-1] / C0

                def km_path():
                    K, M = plan.assemble_km(er_m, ur_m, prec)
                    return cm.to_csc_combine(K, M, k0, block_symmetric=True)
                km_path()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(3):
                    ref = km_path()
                torch.cuda.synchronize()
                km["ms_km_path_without_prep"] = (time.perf_counter() - t0) * 1e3 / 3
                km["ref"] = ref
                del va, er_m, ur_m
                torch.cuda.empty_cache()
            va = VolumeAssembler(f, prec)
            ts = []
            for i, fr in enumerate(freqs):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                r = va.assemble_freq_matrix(er, ur, tand, cond, float(fr))
                torch.cuda.synchronize()
                if i:
                    ts.append((time.perf_counter() - t0) * 1e3)
            rec = {"ms_per_frequency": float(np.median(ts)), "ms_each_after_first": ts,
                   "nnz": int(r.data.numel()),
                   "is_frequency_dependent": r.is_frequency_dependent,
                   "assemblies": va.assemblies}
            if km:
                rec["ms_km_path_without_prep"] = km["ms_km_path_without_prep"]
                rec["rel_diff_vs_km_path"] = float(
                    torch.linalg.norm((r.data - km["ref"]).to(torch.complex128)) /
                    torch.linalg.norm(km["ref"].to(torch.complex128)))
                del km["ref"]
            res[f"{prec}_{name}"] = rec
            log(f"volume step {prec} {name}: {rec}")
            del va, r
            torch.cuda.empty_cache()
    return res


def suite_heat(args) -> dict:
    """Lagrange-2 heat assembly (conductance and capacity, FP64 COO) on a ~2.8M-tet mesh
    (the bench mesh), its CSC map and the per-step to_csc(K + M / dt); CPU line: the
    oracle's restatement of grad.py's builders on a bounded sample (16 threads)."""
    import numpy as np
    import torch
    from oracle import heat as oh
    from emerge_amd.heat import HeatCscMap, HeatPlan
    from emerge_amd.mesh import Legrange2Field, delaunay_box_mesh
    f = Legrange2Field(delaunay_box_mesh(args.points, seed=args.seed))
    plan = HeatPlan.from_field(f)
    T, Nn = f.n_tets, f.mesh.n_nodes
    rng = np.random.default_rng(args.seed + 9)
    cond = torch.as_tensor(rng.uniform(1, 400, (T, 3)), device="cuda")
    rho = torch.as_tensor(rng.uniform(1e6, 4e6, T), device="cuda")
    K, M = plan.assemble(cond=cond, rho_cp=rho)
    ms, _ = time_launches(lambda: plan.assemble(cond=cond, rho_cp=rho), 10, 2)
    # per tet: 16 ids + 24 Nn/T coordinates + 24 (diagonal conductivity) + 8 (rho) in,
    # 2 x 800 out
    b = 16 + 24 * Nn / T + 24 + 8 + 1600
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cm = HeatCscMap(plan)
    torch.cuda.synchronize()
    ms_map = (time.perf_counter() - t0) * 1e3
    # diagonal conductivity: every block is bit-symmetric, so the mirrored reads apply
    ms_csc, _ = time_launches(lambda: cm.to_csc(K, M, 1e3, block_symmetric=True), 10, 2)
    ms_csc_plain, _ = time_launches(lambda: cm.to_csc(K, M, 1e3), 10, 2)
    ms_csc_seg, _ = time_launches(lambda: cm.to_csc(K, M, 1e3, segment_gather=True), 10, 2)
    ref = cm.to_csc(K, M, 1e3, segment_gather=True)
    same = bool(torch.equal(cm.to_csc(K, M, 1e3), ref)) and \
        bool(torch.equal(cm.to_csc(K, M, 1e3, block_symmetric=True), ref))
    n = min(T, 100_000)
    m = f.mesh
    C = np.zeros((3, 3, n))
    cnp = cond[:n].cpu().numpy()
    for d in range(3):
        C[d, d] = cnp[:, d]
    t0 = time.perf_counter()
    oh.build(m.nodes, m.tets[:, :n], f.tet_to_field[:, :n], cond=C, rho_cp=rho[:n].cpu().numpy())
    cpu_tps = n / (time.perf_counter() - t0)
    del K, M, cm
    torch.cuda.empty_cache()
    return {"n_tets": T, "n_dofs": f.n_field, "ms_assemble_K_and_M": ms,
            "tets_per_s": T / (ms * 1e-3), "bytes_per_tet": b,
            "gb_per_s": T * b / (ms * 1e-3) / 1e9,
            "roofline_frac": T * b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "ms_csc_map_one_time": ms_map, "ms_to_csc_K_plus_M_over_dt": ms_csc,
            "ms_to_csc_K_plus_M_over_dt_plain_incidence": ms_csc_plain,
            "ms_to_csc_K_plus_M_over_dt_segment_gather": ms_csc_seg,
            "incidence_equals_segment_gather": same,
            "cpu_oracle_tets_per_s": cpu_tps,
            "cpu_sample": f"first {n} tets, K and M, OpenMP threads"}


def target_run(args) -> dict:
    """The large single-mesh case: mixed and FP64 fused A on a ~8M-tet Delaunay mesh."""
    import numpy as np
    import torch
    from emerge_amd import autotune
    from emerge_amd.assembly import Ned2Plan
    from emerge_amd.mesh import Nedelec2Field, delaunay_box_mesh
    f = Nedelec2Field(delaunay_box_mesh(args.target_points, seed=args.seed + 6))
    plan = Ned2Plan.from_field(f)
    T, Nn = plan.n_tets, f.mesh.n_nodes
    eps, mu = _mats(T, np.random.default_rng(args.seed + 6))
    er, ur = plan.material(eps), plan.material(mu)
    k0 = 2 * np.pi * 10e9 / C0
    log(f"target: {T} tets")
    res = {"n_tets": T, "points": args.target_points}
    a32 = torch.empty(plan.nnz_coo, dtype=torch.complex64, device="cuda")
    t32 = autotune.tune(plan, er, ur, k0, "mixed", out=a32[None], reps=3)
    ms32, _ = time_launches(lambda: plan.assemble_a(er, ur, k0, "mixed", out=a32), 10, 2)
    a64 = torch.empty(plan.nnz_coo, dtype=torch.complex128, device="cuda")
    t64 = autotune.tune(plan, er, ur, k0, "fp64", out=a64[None], reps=3)
    ms64, _ = time_launches(lambda: plan.assemble_a(er, ur, k0, "fp64", out=a64), 10, 2)
    res["tuning"] = {"mixed": tuning_record(t32), "fp64": tuning_record(t64)}
    autotune.tune(plan, er, ur, k0, "mixed", out=a32[None], reps=3)
    plan.assemble_a(er, ur, k0, "mixed", out=a32)
    torch.cuda.synchronize()
    err = rel_err_chunked(a32, a64)
    for name, ms, c in (("mixed", ms32, 8), ("fp64", ms64, 16)):
        b = bytes_per_tet(c, Nn / T)
        res[name] = {"ms_per_launch": ms, "tets_per_s": T / (ms * 1e-3),
                     "gb_per_s": T * b / (ms * 1e-3) / 1e9,
                     "roofline_frac": T * b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    res["rel_err_mixed_vs_fp64"] = err
    del a32, a64
    torch.cuda.empty_cache()
    return res


def load_traffic(workload_key: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE) and
    calibrated on a copy kernel; committed under profiles/ by the profiling recipe."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as fh:
        data = json.load(fh)
    entry = data.get(workload_key)
    return None if entry is None else entry.get("bytes_per_launch")


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--points", type=int, default=400_000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-chunk", type=int, default=50_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cold", action="store_true",
                    help="skip the launches after a 1 GiB fill (cold memory-side cache)")
    ap.add_argument("--no-csc", action="store_true",
                    help="skip the per-frequency CSC line (CSC map, K/M sums, A_f combine)")
    ap.add_argument("--suite", action="store_true",
                    help="also measure the frequency sweep (500k tets x 64 f), the uneven "
                         "32-mesh batch and the backward on 1M tets")
    ap.add_argument("--only", default="",
                    help="with --suite: comma list among sweep,batch,backward,csc_pec,"
                         "surface_interp,heat,materials,ref,dropin")
    ap.add_argument("--target-points", type=int, default=0,
                    help="also run the large single-mesh case with this many points "
                         "(1200000 gives about 8M tets)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain 2 GiB device copy, so that a rocprofv3 --pmc "
                         "run can scale FETCH_SIZE/WRITE_SIZE by a kernel of known traffic")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the assembly shards by tet range with no exchange "
              "between GPUs, and only the single-GPU figure is reported here "
              f"(asked for --gpus {args.gpus}, WORLD_SIZE {world}).", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from emerge_amd.assembly import Ned2Plan
    from emerge_amd.mesh import Nedelec2Field, delaunay_box_mesh

    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    mesh = delaunay_box_mesh(args.points, seed=args.seed)
    field = Nedelec2Field(mesh)
    T, Nn = field.n_tets, mesh.n_nodes
    mesh_s = time.perf_counter() - t0
    log(f"mesh: {T} tets, {Nn} nodes, {field.n_field} DOFs in {mesh_s:.1f}s")

    rng = np.random.default_rng(args.seed + 1)
    eps_np = rng.uniform(1, 10, T) * (1 - 1j * rng.uniform(0, 0.02, T))
    mu_np = rng.uniform(1, 2, (T, 3)).astype(np.complex128)
    plan = Ned2Plan.from_field(field)
    eps = plan.material(torch.from_numpy(eps_np))
    mu = plan.material(torch.from_numpy(mu_np))
    k0 = 2 * np.pi * 10e9 / C0
    nnz = plan.nnz_coo
    log(f"plan ready ({time.perf_counter() - t0:.1f}s); COO entries {nnz}")

    from emerge_amd import autotune
    a32 = torch.empty(nnz, dtype=torch.complex64, device="cuda")
    r = autotune.tune(plan, eps, mu, k0, "mixed", out=a32[None])
    tuning = {"mixed": tuning_record(r)}
    ms32, wall32 = time_launches(lambda: plan.assemble_a(eps, mu, k0, "mixed", out=a32),
                                 args.steps, args.warmup)
    log(f"mixed: {ms32:.3f} ms/launch")
    # the same launch with its inputs out of the memory-side cache, at the tuned pair and
    # at windows of 512 / 1,024 / 4,096 tets (128-tet tiles)
    cold_rec = None
    if not args.no_cold:
        b32c = bytes_per_tet(8, Nn / T)
        cold = {"tuned": cold_cache_launches(
            lambda: plan.assemble_a(eps, mu, k0, "mixed", out=a32))}
        for w in (512, 1024, 4096):
            plan.set_window(w)
            plan.set_tuning(128)
            for pf in (0, 1 << 20):
                plan.set_prefetch(pf)
                cold[f"{w}/128" + ("/pf" if pf else "")] = cold_cache_launches(
                    lambda: plan.assemble_a(eps, mu, k0, "mixed", out=a32))
        apply_tuning(plan, tuning["mixed"])
        cold_best = min(cold.values())
        cold_rec = {"ms_per_launch": cold, "best_ms": cold_best,
                    "roofline_frac_best": T * b32c / (cold_best * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "note": "each launch after a 1 GiB device fill (median of 8); '/pf': "
                            "prefetched launches (chunks of 2^20 tets, each after a kernel "
                            "reading its inputs); the headline times back-to-back launches, "
                            "whose inputs (~240 MB) stay in the 256 MB memory-side cache"}
        log(f"mixed, cold cache: {cold}")
    res64 = None
    err = None
    if not args.no_fp64:
        a64 = torch.empty(nnz, dtype=torch.complex128, device="cuda")
        r = autotune.tune(plan, eps, mu, k0, "fp64", out=a64[None])
        tuning["fp64"] = tuning_record(r)
        ms64, wall64 = time_launches(lambda: plan.assemble_a(eps, mu, k0, "fp64", out=a64),
                                     args.steps, args.warmup)
        log(f"fp64: {ms64:.3f} ms/launch")
        apply_tuning(plan, tuning["mixed"])
        plan.assemble_a(eps, mu, k0, "mixed", out=a32)
        torch.cuda.synchronize()
        err = rel_err_chunked(a32, a64)
        log(f"mixed vs fp64 relative Frobenius error {err:.3e}")
        b64 = bytes_per_tet(16, Nn / T)
        res64 = {"tets_per_s": T / (ms64 * 1e-3), "ms_per_step": ms64,
                 "gb_per_s": T * b64 / (ms64 * 1e-3) / 1e9,
                 "roofline_frac": T * b64 / (ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                 "bytes_per_tet": b64, "dtype": "f64", "output": "complex128 COO"}
        del a64

    if args.copy_calibration:
        src = torch.ones(1 << 31, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("copy calibration: 2 GiB device copy done")

    csc = None
    if not args.no_csc:
        csc = csc_per_frequency(plan, eps, mu, k0, args.steps, args.warmup)
        csc["mixed"]["traffic"] = load_traffic(f"delaunay_p{args.points}_s{args.seed}_csc_freq_mixed")

    b32 = bytes_per_tet(8, Nn / T)
    achieved = T * b32 / (ms32 * 1e-3) / 1e9
    workload_key = f"delaunay_p{args.points}_s{args.seed}_mixed"
    traffic = load_traffic(workload_key)

    cpu = None
    if not args.no_cpu:
        er_ref = np.zeros((3, 3, T), np.complex128)
        ur_ref = np.zeros((3, 3, T), np.complex128)
        for d in range(3):
            er_ref[d, d] = eps_np
            ur_ref[d, d] = mu_np[:, d]
        host = host_cpu_info()
        log(f"cpu baseline (oracle) ... host {host}")
        tps, nth, done, el = cpu_baseline(field, er_ref, ur_ref, args.cpu_seconds, args.cpu_chunk,
                                          nthreads=host["job_threads"])
        cpu = {"value": tps, "unit": "tets/s", "cores": nth, "threads": nth, "kind": "port",
               "sample": f"first {min(args.cpu_chunk, T)} tets of the same mesh, repeated "
                         f"({done} tets in {el:.1f}s); K and M complex128 + int64 rows/cols "
                         f"as the reference's _matrix_builder",
               "host_physical_cores": host["physical_cores"], "host_nproc": host["nproc"],
               "host_affinity_cpus": host["affinity_cpus"],
               "threads_note": "one OpenMP thread per core used; the job's share of the host "
                               "is at most 16 threads per GPU"}
        log(f"cpu: {tps:.0f} tets/s on {nth} threads")
        # the reference's default thread count (NUMBA_NUM_THREADS=4)
        tps4, _, done4, el4 = cpu_baseline(field, er_ref, ur_ref, args.cpu_seconds / 2,
                                           args.cpu_chunk, nthreads=4)
        cpu["value_4_threads"] = tps4
        log(f"cpu: {tps4:.0f} tets/s on 4 threads")

    out = {
        "metric": METRIC,
        "value": T / (ms32 * 1e-3),
        "unit": "tets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "Delaunay unit box, 400k seeded uniform points + jittered "
                               "boundary grid; A = K - k0^2 M COO complex64, f = 10 GHz; "
                               "per-tet complex scalar eps_r, diagonal mu_r",
                   "n_tets": T, "n_nodes": Nn, "n_dofs": field.n_field, "nnz_coo": nnz,
                   "points": args.points, "seed": args.seed,
                   "precision": "FP64 geometry, FP32 contraction, complex64 output"},
        "hbm_gb_per_s": achieved,
        "rel_err_vs_fp64": err,
        "wall_s_timed_region": wall32,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_tet": b32, "kernel": "ned2_assemble_kernel<float, MODE_A>"},
        "fp64": res64,
        "tuning": tuning,
        "cold_cache": cold_rec,
        "csc_per_frequency": csc,
        "cpu_baseline": cpu,
    }
    if args.suite or args.target_points:
        del a32
        plan.close()
        torch.cuda.empty_cache()
        extra = {}
        if args.suite:
            runs = (("sweep", "sweep_500k_x64", suite_sweep),
                    ("batch", "batch_32_meshes", suite_batch),
                    ("backward", "backward_1M", suite_backward),
                    ("csc_pec", "csc_pec_500k", suite_csc_pec),
                    ("surface_interp", "surface_interp_500k", suite_surface_interp),
                    ("heat", "heat_2.8M", suite_heat),
                    ("materials", "materials_2.8M", suite_materials),
                    ("volume", "volume_step_2.8M", suite_volume_step),
                    ("ref", "fp64_ref_500k", suite_ref),
                    ("dropin", "dropin_500k", suite_dropin))
            only = set(args.only.split(",")) if args.only else None
            for short, key, fn in runs:
                if only is None or short in only:
                    log(f"suite: {short}")
                    extra[key] = fn(args)
        if args.target_points:
            extra["target"] = target_run(args)
        out["configs"] = extra
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
