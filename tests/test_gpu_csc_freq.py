"""Per-frequency CSC values from CSC sums of K and M (em_csc_freq_combine, FreqCsc,
CscSweepGraph) against the oracle's to_csc(Evec - Bvec * K0**2) (assembler.py:494,
csc_cast.py:55-63, 162-177): FP64 within 1e-12 relative Frobenius, mixed within 1e-6 of
the FP64 oracle; the pattern is the map's, bit-exact against the oracle."""
import numpy as np
import pytest

import oracle
from tests.conftest import ref_tensor, rel_fro

pytestmark = pytest.mark.gpu

C0 = 299792458.0
K0S = 2 * np.pi * np.linspace(8e9, 12e9, 5) / C0


def _oracle_csc(field, er, ur):
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(field, er, ur, np.array([]),
                                                            local_frame=True)
    return dE, dB, oracle.CSCMap(rows, cols, field.n_field)


@pytest.mark.parametrize("kinds", [("scalar", "diag"), ("full", "sym")])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_freq_csc_matches_oracle(delaunay_field, kinds, precision):
    from emerge_amd.assembly import CscMap, FreqCsc, Ned2Plan
    f = delaunay_field
    rng = np.random.default_rng(40)
    er, ur = ref_tensor(f.n_tets, kinds[0], rng), ref_tensor(f.n_tets, kinds[1], rng)
    dE, dB, omap = _oracle_csc(f, er, ur)
    plan = Ned2Plan.from_field(f)
    cm = CscMap(plan)
    ex = cm.export("indptr", "indices")
    assert np.array_equal(ex["indptr"].cpu().numpy(), omap.indptr)
    assert np.array_equal(ex["indices"].cpu().numpy(), omap.indices)
    K, M = plan.assemble_km(er, ur, precision)
    fc = FreqCsc(cm, K, M, block_symmetric=kinds[0] != "full")
    tol = 1e-12 if precision == "fp64" else 1e-6
    Af = fc.A(K0S).cpu().numpy()
    for i, k0 in enumerate(K0S):
        ref = omap.to_csc_data(dE - dB * k0 ** 2)
        assert rel_fro(Af[i].astype(np.complex128), ref) <= tol
        assert rel_fro(fc.A(float(k0)).cpu().numpy().astype(np.complex128), ref) <= tol


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_csc_freq_combine_edges(dtype):
    """Odd lengths (the complex64 tail entry), more than 128 wavenumbers (two launches), a
    padded output stride, and an empty array."""
    import torch
    from emerge_amd.assembly import csc_freq_combine
    dt = getattr(torch, dtype)
    g = torch.Generator(device="cuda").manual_seed(3)
    for n in (1, 7, 1001, 65537):
        K = torch.randn(n, dtype=dt, device="cuda", generator=g)
        M = torch.randn(n, dtype=dt, device="cuda", generator=g)
        k0 = np.linspace(100.0, 300.0, 130)
        stride = n + (n % 2) + 64
        out = torch.full((130, stride), complex("nan"), dtype=dt, device="cuda")
        csc_freq_combine(K, M, k0, out=out[:, :n])
        Kd, Md = K.to(torch.complex128), M.to(torch.complex128)
        for i in (0, 1, 127, 128, 129):
            ref = Kd - k0[i] ** 2 * Md
            got = out[i, :n].to(torch.complex128)
            assert float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref)) <= \
                (1e-15 if dtype == "complex128" else 1e-7)
        assert torch.isnan(out[:, n:].real).all()  # padding untouched
    empty = torch.empty(0, dtype=dt, device="cuda")
    assert csc_freq_combine(empty, empty, [1.0, 2.0]).shape == (2, 0)
    # a multi-frequency output must have one row per wavenumber
    K = torch.randn(8, dtype=dt, device="cuda", generator=g)
    for bad in (torch.empty((1, 16), dtype=dt, device="cuda"),
                torch.empty((4, 4), dtype=dt, device="cuda"),
                torch.empty((2, 2, 4), dtype=dt, device="cuda")):
        with pytest.raises(ValueError):
            csc_freq_combine(K, K, [1.0, 2.0], out=bad)


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_csc_freq_combine_lossy(dtype):
    """K - k0^2 M + 1j g C (em_csc_freq_combine_lossy) against the same expression in
    torch (FP64), over more than 128 wavenumbers and an odd length."""
    import torch
    from emerge_amd.assembly import csc_freq_combine
    dt = getattr(torch, dtype)
    g_ = torch.Generator(device="cuda").manual_seed(5)
    n = 4097
    K, M, C = (torch.randn(n, dtype=dt, device="cuda", generator=g_) for _ in range(3))
    k0 = np.linspace(100.0, 300.0, 131)
    g = k0 ** 2 / (2 * np.pi * 1e9 * k0 / 100.0 * 8.854187818814e-12)
    out = csc_freq_combine(K, M, k0, C=C, g=g)
    Kd, Md, Cd = (x.to(torch.complex128) for x in (K, M, C))
    for i in (0, 1, 127, 128, 130):
        ref = Kd - k0[i] ** 2 * Md + 1j * g[i] * Cd
        got = out[i].to(torch.complex128)
        assert float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref)) <= \
            (1e-15 if dtype == "complex128" else 1e-7)
    # without C it is the plain combine, bit for bit
    assert torch.equal(csc_freq_combine(K, M, k0[:3]), csc_freq_combine(K, M, k0[:3]))
    with pytest.raises(ValueError):
        csc_freq_combine(K, M, k0, C=C)  # g missing


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_csc_sweep_graph(wr90_field, precision):
    """The whole CSC sweep (K, M, their CSC sums, every A_f) recorded once and replayed
    three times, with eager work and synchronisations between replays: every replay equals
    the eager FreqCsc result bit for bit and the oracle within the tolerance."""
    import torch
    from emerge_amd.assembly import CscMap, FreqCsc, Ned2Plan
    from emerge_amd.graph import CscSweepGraph
    f = wr90_field
    rng = np.random.default_rng(41)
    er, ur = ref_tensor(f.n_tets, "scalar", rng), ref_tensor(f.n_tets, "diag", rng)
    dE, dB, omap = _oracle_csc(f, er, ur)
    plan = Ned2Plan.from_field(f)
    cm = CscMap(plan)
    k0s = 2 * np.pi * np.linspace(8e9, 12e9, 64) / C0
    sg = CscSweepGraph(plan, cm, er, ur, k0s, precision)
    K, M = plan.assemble_km(er, ur, precision)
    eager = FreqCsc(cm, K, M, block_symmetric=True).A(k0s)
    tol = 1e-12 if precision == "fp64" else 1e-6
    for rep in range(3):
        sg.out.zero_()
        torch.cuda.synchronize()
        out = sg.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), rep
        _ = plan.assemble_a(er, ur, float(k0s[rep]), precision)  # eager work in between
        torch.cuda.synchronize()
    for i in (0, 17, 40, 63):
        ref = omap.to_csc_data(dE - dB * k0s[i] ** 2)
        assert rel_fro(sg.out[i].cpu().numpy().astype(np.complex128), ref) <= tol
