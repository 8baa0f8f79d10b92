"""GPU backward kernel (dL/d eps_r, dL/d mu_r^-1 per tet) for L = Re(lam^H A x).

Checked against (1) an FP64 reference gradient built from the oracle: A is linear in
eps_r and mu_r^-1, so dL/d eps_e = -k0^2 sum_k Mhat_k G_k and dL/d mu^-1_{e,d} =
sum_k (K(I + e_d e_d^T) - K(I))_k G_k with G_k = lam[row_k] conj(x[col_k]) and the unit-
material blocks from the oracle; (2) central finite differences of L through the GPU
forward (h = 1e-6, 256 sampled tets, real and imaginary parts)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

C0 = 299792458.0


def _ref_tensor_diag(T, diag):
    out = np.zeros((3, 3, T), complex)
    for d in range(3):
        out[d, d] = diag[d]
    return out


def _setup(field):
    import torch
    from emerge_amd.assembly import Ned2Plan
    plan = Ned2Plan.from_field(field)
    T, N = field.n_tets, field.n_field
    rng = np.random.default_rng(17)
    eps = torch.tensor(rng.uniform(1, 10, T) * (1 - 1j * rng.uniform(0, 0.02, T)),
                       dtype=torch.complex128, device="cuda")
    muinv = torch.tensor(1.0 / rng.uniform(1, 2, (T, 3)), dtype=torch.complex128, device="cuda")
    lam = torch.tensor(rng.normal(size=N) + 1j * rng.normal(size=N), device="cuda")
    x = torch.tensor(rng.normal(size=N) + 1j * rng.normal(size=N), device="cuda")
    rows, cols = plan.rowcol()
    k0 = 2 * np.pi * 10e9 / C0
    return dict(field=field, plan=plan, eps=eps, muinv=muinv, lam=lam, x=x, rows=rows,
                cols=cols, k0=k0, T=T, rng=rng)


@pytest.fixture(scope="module")
def setup():
    from emerge_amd.mesh import Nedelec2Field, delaunay_box_mesh
    return _setup(Nedelec2Field(delaunay_box_mesh(3000, seed=17)))


def _grad(s, precision="fp64"):
    from emerge_amd.autograd import assemble_a, bilinear_form
    e = s["eps"].clone().requires_grad_(True)
    m = s["muinv"].clone().requires_grad_(True)
    A = assemble_a(s["plan"], e, m, s["k0"], precision)
    L = bilinear_form(A, s["rows"], s["cols"], s["lam"], s["x"])
    L.backward()
    return e.grad.cpu().numpy(), m.grad.cpu().numpy(), float(L.detach())


def _check_reference_gradient(s):
    f, T, k0 = s["field"], s["T"], s["k0"]
    ge, gm, _ = _grad(s)
    I = _ref_tensor_diag(T, [1, 1, 1])
    _, Mhat, rows, cols = oracle.tet_mass_stiffness_matrices(f, I, I, [], local_frame=True)
    K0 = oracle.tet_mass_stiffness_matrices(f, I, I, [], local_frame=True)[0]
    lam, x = s["lam"].cpu().numpy(), s["x"].cpu().numpy()
    G = lam[rows] * np.conj(x[cols])
    ref_e = -k0 ** 2 * (Mhat * G).reshape(T, 400).sum(axis=1)
    assert np.linalg.norm(ge - ref_e) <= 1e-12 * np.linalg.norm(ref_e)
    for d in range(3):
        mu = [1.0, 1.0, 1.0]
        mu[d] = 0.5  # mu^-1 = I + e_d e_d^T
        Kd = oracle.tet_mass_stiffness_matrices(f, I, _ref_tensor_diag(T, mu), [],
                                                local_frame=True)[0]
        ref_m = ((Kd - K0) * G).reshape(T, 400).sum(axis=1)
        assert np.linalg.norm(gm[:, d] - ref_m) <= 1e-10 * np.linalg.norm(ref_m), d


def test_backward_matches_fp64_reference_gradient(setup):
    _check_reference_gradient(setup)


@pytest.mark.parametrize("order", [(0, 1, 2, 3), (1, 0, 2, 3)])
def test_backward_one_tet(order):
    """A one-tet plan in both orientations: one wavefront tile holding a single tet."""
    from emerge_amd.mesh import Nedelec2Field, TetMesh
    nodes = np.array([[0, 1.0, 0, 0], [0, 0, 1.0, 0], [0, 0, 0, 1.0]]) * 1e-2
    field = Nedelec2Field(TetMesh(nodes, np.array(order, np.int64).reshape(4, 1)))
    _check_reference_gradient(_setup(field))


def test_backward_matches_central_differences(setup):
    import torch
    from emerge_amd.autograd import bilinear_form
    s = setup
    plan, k0, T = s["plan"], s["k0"], s["T"]
    ge, gm, _ = _grad(s)
    h = 1e-6
    sample = s["rng"].choice(T, size=256, replace=False)

    def Lt(eps, muinv):
        # L = sum over tets of these per-tet terms; the difference of two evaluations is
        # taken per tet before summing, so the unperturbed tets cancel exactly
        A = plan.assemble_a(plan.material(eps), plan.material(muinv, inverse=True), k0, "fp64")
        w = torch.conj(s["lam"][s["rows"]]) * A * s["x"][s["cols"]]
        return torch.real(w).reshape(T, 400).sum(dim=1)

    def dL(a, b):
        return float(torch.sum(a - b))

    fd_e = np.zeros(256, complex)
    fd_m = np.zeros(256, complex)
    for n, t in enumerate(sample):
        for part, unit in ((0, 1.0), (1, 1j)):
            e1, e2 = s["eps"].clone(), s["eps"].clone()
            e1[t] += unit * h
            e2[t] -= unit * h
            d = dL(Lt(e1, s["muinv"]), Lt(e2, s["muinv"])) / (2 * h)
            fd_e[n] += d if part == 0 else 1j * d
            m1, m2 = s["muinv"].clone(), s["muinv"].clone()
            m1[t, 1] += unit * h
            m2[t, 1] -= unit * h
            d = dL(Lt(s["eps"], m1), Lt(s["eps"], m2)) / (2 * h)
            fd_m[n] += d if part == 0 else 1j * d
    scale_e = np.abs(ge).max()
    scale_m = np.abs(gm[:, 1]).max()
    assert np.max(np.abs(fd_e - ge[sample])) <= 1e-5 * scale_e
    assert np.max(np.abs(fd_m - gm[sample, 1])) <= 1e-5 * scale_m


def test_full_tensor_leaves_and_mixed_forward(setup):
    """(T,3,3) leaves (antisymmetric terms included) against the reference gradient built
    by linearity from the FP64 forward; the mixed forward gives the same FP64 gradient."""
    import torch
    from emerge_amd.autograd import assemble_a, bilinear_form
    s = setup
    plan, k0, T = s["plan"], s["k0"], s["T"]
    rng = np.random.default_rng(5)
    E = torch.tensor(rng.normal(size=(T, 3, 3)) * 0.2 + 3 * np.eye(3) + 0.1j, device="cuda")
    Mi = torch.tensor(rng.normal(size=(T, 3, 3)) * 0.1 + np.eye(3), device="cuda",
                      dtype=torch.complex128)
    e = E.clone().requires_grad_(True)
    m = Mi.clone().requires_grad_(True)
    L = bilinear_form(assemble_a(plan, e, m, k0), s["rows"], s["cols"], s["lam"], s["x"])
    L.backward()
    G = s["lam"][s["rows"]] * torch.conj(s["x"][s["cols"]])
    for comp in (0, 1, 5, 7):
        p, q = divmod(comp, 3)
        unit = torch.zeros((T, 3, 3), dtype=torch.complex128, device="cuda")
        unit[:, p, q] = 1.0
        zero = torch.zeros_like(unit)
        dM = plan.assemble_a(plan.material(unit), plan.material(zero, inverse=True), k0, "fp64")
        ref_e = (dM * G).reshape(T, 400).sum(dim=1)
        dK = plan.assemble_a(plan.material(zero), plan.material(unit, inverse=True), k0, "fp64")
        ref_m = (dK * G).reshape(T, 400).sum(dim=1)
        assert torch.linalg.norm(e.grad[:, p, q] - ref_e) <= 1e-11 * torch.linalg.norm(ref_e)
        assert torch.linalg.norm(m.grad[:, p, q] - ref_m) <= 1e-11 * torch.linalg.norm(ref_m)
    # the mixed forward's output is complex64, so autograd hands the backward an upstream
    # gradient rounded to complex64; the gradient itself is still computed in FP64
    ge64, gm64, _ = _grad(s, "fp64")
    ge32, gm32, _ = _grad(s, "mixed")
    assert np.linalg.norm(ge32 - ge64) <= 1e-6 * np.linalg.norm(ge64)
    assert np.linalg.norm(gm32 - gm64) <= 1e-6 * np.linalg.norm(gm64)


def test_subset_plan_scalar_and_diag_leaves(setup):
    """Conductor tets removed (a plan over a subset of 70% of the tets, not a multiple of
    the kernel's 64-tet tiles), scalar and diagonal leaves, one leaf at a time: the
    gradient of every assembled tet equals the linear-term reference and removed tets get
    zero."""
    import torch
    from emerge_amd.assembly import Ned2Plan
    from emerge_amd.autograd import assemble_a, bilinear_form
    s = setup
    f, T, k0 = s["field"], s["T"], s["k0"]
    rng = np.random.default_rng(23)
    keep = np.sort(rng.choice(T, size=int(0.7 * T), replace=False))
    assert len(keep) % 64 != 0
    plan = Ned2Plan.from_field(f, tetids=keep)
    rows, cols = plan.rowcol()
    lam, x = s["lam"], s["x"]
    G = lam[rows] * torch.conj(x[cols])
    ids = torch.as_tensor(keep, device="cuda")
    eps = s["eps"].clone()
    mu = s["muinv"].clone()
    # scalar eps only
    e = eps.clone().requires_grad_(True)
    L = bilinear_form(assemble_a(plan, e, mu, k0), rows, cols, lam, x)
    L.backward()
    one = torch.ones(T, dtype=torch.complex128, device="cuda")
    zero3 = torch.zeros((T, 3), dtype=torch.complex128, device="cuda")
    dM = plan.assemble_a(plan.material(one), plan.material(zero3, inverse=True), k0, "fp64")
    ref = torch.zeros(T, dtype=torch.complex128, device="cuda")
    ref[ids] = (dM * G).reshape(-1, 400).sum(dim=1)
    assert torch.linalg.norm(e.grad - ref) <= 1e-12 * torch.linalg.norm(ref)
    removed = torch.ones(T, dtype=torch.bool, device="cuda")
    removed[ids] = False
    assert torch.all(e.grad[removed] == 0)
    # diagonal mu^-1 only
    m = mu.clone().requires_grad_(True)
    L = bilinear_form(assemble_a(plan, eps, m, k0), rows, cols, lam, x)
    L.backward()
    zero = torch.zeros(T, dtype=torch.complex128, device="cuda")
    for d in range(3):
        unit = torch.zeros((T, 3), dtype=torch.complex128, device="cuda")
        unit[:, d] = 1.0
        dK = plan.assemble_a(plan.material(zero), plan.material(unit, inverse=True), k0, "fp64")
        ref = torch.zeros(T, dtype=torch.complex128, device="cuda")
        ref[ids] = (dK * G).reshape(-1, 400).sum(dim=1)
        assert torch.linalg.norm(m.grad[:, d] - ref) <= 1e-12 * torch.linalg.norm(ref), d
        assert torch.all(m.grad[removed, d] == 0)


def test_full_and_diagonal_leaves_together(setup):
    """A full (T,3,3) eps_r leaf with a diagonal mu_r^-1 leaf (the full-tensor kernel writes
    each leaf in its own kind), against the linear-term reference."""
    import torch
    from emerge_amd.autograd import assemble_a, bilinear_form
    s = setup
    plan, k0, T = s["plan"], s["k0"], s["T"]
    rng = np.random.default_rng(9)
    E = torch.tensor(rng.normal(size=(T, 3, 3)) * 0.2 + 3 * np.eye(3) - 0.05j, device="cuda")
    e = E.clone().requires_grad_(True)
    m = s["muinv"].clone().requires_grad_(True)
    L = bilinear_form(assemble_a(plan, e, m, k0), s["rows"], s["cols"], s["lam"], s["x"])
    L.backward()
    G = s["lam"][s["rows"]] * torch.conj(s["x"][s["cols"]])
    zero9 = torch.zeros((T, 3, 3), dtype=torch.complex128, device="cuda")
    zero3 = torch.zeros((T, 3), dtype=torch.complex128, device="cuda")
    for comp in (0, 2, 4, 6):
        p, q = divmod(comp, 3)
        unit = zero9.clone()
        unit[:, p, q] = 1.0
        dM = plan.assemble_a(plan.material(unit), plan.material(zero3, inverse=True), k0, "fp64")
        ref = (dM * G).reshape(T, 400).sum(dim=1)
        assert torch.linalg.norm(e.grad[:, p, q] - ref) <= 1e-11 * torch.linalg.norm(ref)
    for d in range(3):
        unit = zero3.clone()
        unit[:, d] = 1.0
        dK = plan.assemble_a(plan.material(zero9), plan.material(unit, inverse=True), k0, "fp64")
        ref = (dK * G).reshape(T, 400).sum(dim=1)
        assert torch.linalg.norm(m.grad[:, d] - ref) <= 1e-11 * torch.linalg.norm(ref), d
