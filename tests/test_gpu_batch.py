"""GPU tests of the batched launches: tile-size variants, the autotuner, uneven
multi-mesh batches in one launch, and frequency sweeps replayed from a HIP graph."""
import numpy as np
import pytest

import oracle
from tests.conftest import ref_tensor, rel_fro

pytestmark = pytest.mark.gpu

C0 = 299792458.0


def _mats(T, rng):
    import torch
    eps = torch.tensor(rng.uniform(1, 10, T) * (1 - 1j * rng.uniform(0, 0.02, T)),
                       dtype=torch.complex128)
    mu = torch.tensor(rng.uniform(1, 2, (T, 3)), dtype=torch.complex128)
    return eps, mu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tile_variants_identical(delaunay_field, precision):
    import torch
    from emerge_amd.assembly import Ned2Plan
    plan = Ned2Plan.from_field(delaunay_field)
    eps, mu = _mats(plan.n_tets, np.random.default_rng(1))
    k0 = 2 * np.pi * 9e9 / C0
    outs = []
    for tile in (32, 64, 128):
        plan.set_tuning(tile)
        assert plan.get_tuning() == tile
        outs.append(plan.assemble_a(eps, mu, k0, precision).cpu())
        K, M = plan.assemble_km(eps, mu, precision)
        if tile == 32:
            K1, M1 = K.cpu(), M.cpu()
        else:
            assert torch.equal(K.cpu(), K1) and torch.equal(M.cpu(), M1)
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])


def test_autotuner_picks_and_caches(delaunay_field, tmp_path, monkeypatch):
    from emerge_amd import autotune
    from emerge_amd.assembly import Ned2Plan
    monkeypatch.setenv("EMERGE_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    plan = Ned2Plan.from_field(delaunay_field)
    eps, mu = _mats(plan.n_tets, np.random.default_rng(2))
    k0 = 2 * np.pi * 10e9 / C0
    r = autotune.tune(plan, eps, mu, k0, "mixed", reps=2, warmup=1)
    assert r["nb"] in (32, 64, 128) and r["window"] in autotune.WINDOWS and not r["cached"]
    assert set(r["ms"]) == {f"{w}/{nb}" + (f"/pf{pf}" if pf else "") for w in autotune.WINDOWS
                            for nb in (32, 64, 128) for pf in autotune.PREFETCH}
    assert r["prefetch"] in autotune.PREFETCH
    assert plan.get_tuning() == r["nb"] and plan.get_window() == r["window"]
    assert plan.get_prefetch() == r["prefetch"]
    plan2 = Ned2Plan.from_field(delaunay_field)
    assert plan2.get_window() == 1024 and plan2.get_prefetch() == 0
    r2 = autotune.tune(plan2, eps, mu, k0, "mixed")
    assert r2["cached"] and r2["nb"] == r["nb"] and plan2.get_tuning() == r["nb"]
    assert plan2.get_window() == r["window"] and plan2.get_prefetch() == r["prefetch"]


def test_autotuner_cold_cache_mode(delaunay_field, tmp_path, monkeypatch):
    """cold=True times single launches after a device fill, keeps its own cache entry and
    leaves the outputs unchanged."""
    import torch
    from emerge_amd import autotune
    from emerge_amd.assembly import Ned2Plan
    monkeypatch.setenv("EMERGE_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    plan = Ned2Plan.from_field(delaunay_field)
    eps, mu = _mats(plan.n_tets, np.random.default_rng(3))
    k0 = 2 * np.pi * 10e9 / C0
    ref = plan.assemble_a(eps, mu, k0, "mixed").cpu()
    r = autotune.tune(plan, eps, mu, k0, "mixed", reps=2, cold=True, flush_bytes=1 << 24,
                      windows=(512, 1024), candidates=(64, 128), rounds=1,
                      prefetch=(0, 4096))
    assert r["key"].endswith("|cold") and not r["cached"]
    assert set(r["ms"]) == {f"{w}/{nb}" + p for w in (512, 1024) for nb in (64, 128)
                            for p in ("", "/pf4096")}
    assert all(v > 0 for v in r["ms"].values())
    assert plan.get_tuning() == r["nb"] and plan.get_window() == r["window"]
    assert torch.equal(plan.assemble_a(eps, mu, k0, "mixed").cpu(), ref)
    warm = autotune.tune(plan, eps, mu, k0, "mixed", reps=2, windows=(512, 1024),
                         candidates=(64, 128), rounds=1)
    assert not warm["cached"] and warm["key"] != r["key"]
    again = autotune.tune(plan, eps, mu, k0, "mixed", cold=True, windows=(512, 1024),
                          candidates=(64, 128))
    assert again["cached"] and again["nb"] == r["nb"] and again["window"] == r["window"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_window_variants_identical(delaunay_field, precision):
    """The window only reorders the tiles: K, M and A are the same bit for bit at every
    window (including one larger than the mesh and one of 32 tets), and equal the oracle."""
    import torch
    from emerge_amd.assembly import Ned2Plan
    plan = Ned2Plan.from_field(delaunay_field)
    eps, mu = _mats(plan.n_tets, np.random.default_rng(5))
    k0 = 2 * np.pi * 9e9 / C0
    ref = None
    for w in (1024, 32, 512, 4096, 1 << 20):
        plan.set_window(w)
        assert plan.get_window() == w
        for tile in (32, 128):
            plan.set_tuning(tile)
            A = plan.assemble_a(eps, mu, k0, precision).cpu()
            if ref is None:
                ref = A
            assert torch.equal(A, ref), (w, tile)
    with pytest.raises(Exception):
        plan.set_window(16)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_prefetched_launches_identical(delaunay_field, precision):
    """Prefetched launches (the tile list cut into chunks, each launched after a kernel
    that reads its inputs) give K, M, A and a three-wavenumber A equal bit for bit to one
    launch: several chunk sizes (one chunk per 4,096 tets up to one for the whole mesh),
    windows and tiles, per-tet and (3,3,T) full tensors, conductor tets removed."""
    import torch
    from emerge_amd.assembly import Ned2Plan
    T = delaunay_field.n_tets
    rng = np.random.default_rng(9)
    keep = np.sort(rng.choice(T, size=T - T // 7, replace=False))
    k0 = 2 * np.pi * 9e9 / C0
    k0s = [0.7 * k0, k0, 1.3 * k0]
    eps, mu = _mats(T, rng)
    er_full, ur_full = ref_tensor(T, "full", rng), ref_tensor(T, "full", rng)
    for tetids in (None, keep):
        plan = Ned2Plan.from_field(delaunay_field, tetids=tetids)
        assert plan.get_prefetch() == 0
        for er, ur in ((eps, mu), (er_full, ur_full)):
            for w, tile in ((1024, 128), (512, 32)):
                plan.set_window(w)
                plan.set_tuning(tile)
                plan.set_prefetch(0)
                ref = [plan.assemble_a(er, ur, k0, precision).cpu(),
                       *[x.cpu() for x in plan.assemble_km(er, ur, precision)],
                       plan.assemble_a(er, ur, k0s, precision).cpu()]
                for chunk in (4096, 6000, 1 << 20):
                    plan.set_prefetch(chunk)
                    assert plan.get_prefetch() == chunk
                    got = [plan.assemble_a(er, ur, k0, precision).cpu(),
                           *[x.cpu() for x in plan.assemble_km(er, ur, precision)],
                           plan.assemble_a(er, ur, k0s, precision).cpu()]
                    for a, b in zip(got, ref):
                        assert torch.equal(a, b), (tetids is None, w, tile, chunk)
        # the chunks follow a window change
        plan.set_prefetch(4096)
        plan.set_window(2048)
        plan.set_prefetch(0)
        A0 = plan.assemble_a(eps, mu, k0, precision).cpu()
        plan.set_prefetch(4096)
        assert torch.equal(plan.assemble_a(eps, mu, k0, precision).cpu(), A0)
        for bad in (-1, 1, 4095):
            with pytest.raises(Exception):
                plan.set_prefetch(bad)
        assert plan.get_prefetch() == 4096


def test_uneven_multimesh_batch_one_launch():
    """Meshes of uneven size in one plan: every mesh's COO slice equals the oracle's
    assembly of that mesh alone, and its DOF map is the mesh's own plus an offset."""
    import torch
    from emerge_amd.assembly import Ned2Plan
    from emerge_amd.mesh import Nedelec2Field, delaunay_box_mesh, hex_box_mesh
    fields = [Nedelec2Field(delaunay_box_mesh(n, seed=30 + i, box=(1.0, 0.5 + 0.1 * i, 2.0)))
              for i, n in enumerate([40, 700, 90, 2500, 13])]
    fields.insert(2, Nedelec2Field(hex_box_mesh(0.02, 0.01, 0.03, 3, 2, 4)))
    plan = Ned2Plan.from_fields(fields)
    T = plan.n_tets
    rng = np.random.default_rng(3)
    eps, mu = _mats(T, rng)
    k0 = 2 * np.pi * 10e9 / C0
    A = plan.assemble_a(eps, mu, k0, "fp64").cpu().numpy()
    A32 = plan.assemble_a(eps, mu, k0, "mixed").cpu().numpy()
    rows, cols = plan.rowcol()
    rows, cols = rows.cpu().numpy(), cols.cpu().numpy()
    for i, f in enumerate(fields):
        t0, t1 = plan.tet_offsets[i], plan.tet_offsets[i + 1]
        er = np.zeros((3, 3, f.n_tets), complex)
        ur = np.zeros((3, 3, f.n_tets), complex)
        e_np = eps.numpy()[t0:t1]
        m_np = mu.numpy()[t0:t1]
        for d in range(3):
            er[d, d] = e_np
            ur[d, d] = m_np[:, d]
        dE, dB, r, c = oracle.tet_mass_stiffness_matrices(f, er, ur, [], local_frame=True)
        sl = slice(400 * t0, 400 * t1)
        assert rel_fro(A[sl], dE - k0 ** 2 * dB) <= 1e-12, i
        assert rel_fro(A32[sl].astype(complex), dE - k0 ** 2 * dB) <= 1e-6, i
        assert np.array_equal(rows[sl] - plan.dof_offsets[i], r)
        assert np.array_equal(cols[sl] - plan.dof_offsets[i], c)
    # prefetched launches over the batch (chunks spanning mesh boundaries): the same bits
    plan.set_prefetch(4096)
    assert np.array_equal(plan.assemble_a(eps, mu, k0, "fp64").cpu().numpy(), A)
    assert np.array_equal(plan.assemble_a(eps, mu, k0, "mixed").cpu().numpy(), A32)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sweep_graph_replays_match_eager(delaunay_field, precision):
    import torch
    from emerge_amd.assembly import Ned2Plan
    from emerge_amd.graph import SweepGraph
    plan = Ned2Plan.from_field(delaunay_field)
    eps, mu = _mats(plan.n_tets, np.random.default_rng(4))
    k0s = 2 * np.pi * np.linspace(8e9, 12e9, 64) / C0
    eager = plan.assemble_a(eps, mu, k0s, precision).clone()
    sg = SweepGraph(plan, eps, mu, k0s, precision)
    for rep in range(3):
        sg.out.zero_()
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        out = sg.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), rep
        assert float(junk) == float(junk)
    # every frequency equals the single-k0 launch (which folds -k0^2 into the metric
    # tensor before contracting, so the two round differently)
    tol = 1e-14 if precision == "fp64" else 1e-6
    for f in (0, 31, 63):
        one = plan.assemble_a(eps, mu, float(k0s[f]), precision)
        assert float(torch.linalg.norm(one - eager[f]) / torch.linalg.norm(one)) <= tol
    # with prefetched launches the captured graph holds a prefetch and an assembly launch
    # per chunk; its replays equal the single-launch sweep bit for bit
    plan.set_prefetch(4096)
    sgp = SweepGraph(plan, eps, mu, k0s, precision)
    for rep in range(3):
        sgp.out.zero_()
        junk = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        assert torch.equal(sgp.replay(), eager), rep
        torch.cuda.synchronize()
        assert float(junk) == float(junk)


def test_sweep_more_than_128_frequencies(delaunay_field):
    """Sweeps longer than one launch's wavenumber pack are split into several launches."""
    import torch
    from emerge_amd.assembly import Ned2Plan
    from emerge_amd.mesh import Nedelec2Field, delaunay_box_mesh
    f = Nedelec2Field(delaunay_box_mesh(120, seed=7))
    plan = Ned2Plan.from_field(f)
    eps, mu = _mats(plan.n_tets, np.random.default_rng(5))
    k0s = 2 * np.pi * np.linspace(1e9, 20e9, 300) / C0
    out = plan.assemble_a(eps, mu, k0s, "fp64")
    for fi in (0, 127, 128, 255, 256, 299):
        one = plan.assemble_a(eps, mu, float(k0s[fi]), "fp64")
        assert float(torch.linalg.norm(out[fi] - one) / torch.linalg.norm(one)) <= 1e-14
    # prefetched launches inside every wavenumber pack: the same values bit for bit
    plan.set_prefetch(4096)
    assert torch.equal(plan.assemble_a(eps, mu, k0s, "fp64"), out)
