"""GPU parity of point location and field interpolation (interp.hip, emerge_amd.interp)
against the oracle restatements of index_interp / ned2_tet_interp[_curl], and an end-to-end
two-port WR-90 waveguide: GPU assembly + PEC elimination + Robin ports + GPU interpolation
for the S-parameters."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

from oracle import interp as oi
from tests.conftest import rel_fro

pytestmark = pytest.mark.gpu

C0 = 299792458.0
A_WG, B_WG, L_WG = 22.86e-3, 10.16e-3, 50e-3


def _points(field, rng, n):
    m = field.mesh
    t = rng.integers(0, field.n_tets, n)
    lam = rng.dirichlet(np.ones(4), n).T
    pts = sum(m.nodes[:, m.tets[k, t]] * lam[k] for k in range(4))
    # plus mesh vertices (several containing tets: the largest id wins), face centroids and
    # points outside the box
    v = m.nodes[:, rng.choice(m.n_nodes, 50, replace=False)]
    fc = m.nodes[:, m.tris[:, rng.choice(m.n_tris, 50, replace=False)]].mean(axis=1)
    out = np.array([[1.5, -0.2, 0.5], [0.5, 0.5, 2.0], [0.3, 0.4, 0.5]])
    return np.concatenate([pts, v, fc, out], axis=1)


def test_locate_matches_oracle(delaunay_field):
    import torch
    from emerge_amd.assembly import Ned2Plan
    from emerge_amd.interp import locate
    f = delaunay_field
    plan = Ned2Plan.from_field(f)
    pts = _points(f, np.random.default_rng(0), 20000)
    c = torch.as_tensor(pts, device="cuda")
    got = locate(plan, c).cpu().numpy()
    assert np.array_equal(got, oi.locate(f.mesh, pts))
    sub = np.sort(np.random.default_rng(1).choice(f.n_tets, f.n_tets // 3, replace=False))
    got = locate(plan, c, tetids=sub).cpu().numpy()
    assert np.array_equal(got, oi.locate(f.mesh, pts, tetids=sub))


@pytest.mark.parametrize("curl", [False, True])
def test_interp_matches_oracle(delaunay_field, curl):
    import torch
    from emerge_amd.assembly import Ned2Plan
    from emerge_amd.interp import interp, locate
    f = delaunay_field
    plan = Ned2Plan.from_field(f)
    rng = np.random.default_rng(2)
    pts = _points(f, rng, 5000)
    c = torch.as_tensor(pts, device="cuda")
    tet = locate(plan, c)
    sol = rng.normal(size=f.n_field) + 1j * rng.normal(size=f.n_field)
    ct = rng.normal(size=f.n_tets) + 1j * rng.normal(size=f.n_tets) if curl else None
    got = interp(plan, c, torch.as_tensor(sol), tet, curl=curl,
                 c_tet=None if ct is None else torch.as_tensor(ct)).cpu().numpy()
    ref = oi.interp(f, pts, sol, tet.cpu().numpy(), curl=curl, c_tet=ct, local_frame=True)
    ok = ~np.isnan(ref[0])
    assert np.array_equal(np.isnan(got[0]), ~ok) and ok.sum() > 5000
    assert rel_fro(got[:, ok], ref[:, ok]) <= 1e-12


def test_whitney_field_on_gpu(wr90_field):
    import torch
    from emerge_amd.interp import interpolate_Ef
    f = wr90_field
    E0, b = np.array([1.0, 2.0, -0.5]), np.array([10.0, -5.0, 3.0])
    sol = oi.whitney_coefficients(f, E0, b)
    rng = np.random.default_rng(3)
    xs, ys, zs = (rng.uniform(0, d, 3000) for d in (A_WG, B_WG, L_WG))
    E = interpolate_Ef(f, sol)(xs, ys, zs)
    ref = E0[:, None] + np.cross(b[None, :], np.stack([xs, ys, zs]).T).T
    assert np.abs(E - ref).max() <= 1e-10


def _port_tris(field, z):
    m = field.mesh
    bt = m.boundary_tris()
    zc = m.nodes[2][m.tris[:, bt]]
    return bt[np.all(np.abs(zc - z) < 1e-9, axis=0)]


def test_two_port_waveguide_s_parameters(wr90_field):
    """22.86 x 10.16 x 50 mm WR-90 section, PEC walls, TE10 ports at both ends, 10 GHz:
    the system is assembled, reduced and completed with the port blocks on the GPU, solved
    with SciPy, and the S-parameters come from the GPU interpolation of the field:
    |S11| ~ 0, |S21| ~ 1, arg S21 ~ -beta L (beta = 158.238 rad/m)."""
    import torch
    from emerge_amd.assembly import CscMap, Ned2Plan
    from emerge_amd.interp import interpolate_Ef
    from emerge_amd.pec import pec_solve_ids, scatter_solution, to_scipy
    from emerge_amd.ports import LocalFrame, RectangularWaveguideMode
    from emerge_amd.robin import DPTS, RobinSurface, generate_points_3d
    from emerge_amd.sparam import s_parameter
    f = wr90_field
    k0 = 2 * np.pi * 10e9 / C0
    plan = Ned2Plan.from_field(f)
    p1, p2 = _port_tris(f, 0.0), _port_tris(f, L_WG)
    walls = np.setdiff1d(f.mesh.boundary_tris(), np.concatenate([p1, p2]))
    modes = [RectangularWaveguideMode(LocalFrame([A_WG / 2, B_WG / 2, z], [1, 0, 0], [0, 1, 0],
                                                 [0, 0, 1]), A_WG, B_WG, (1, 0))
             for z in (0.0, L_WG)]
    beta = modes[0].get_beta(k0)
    K, M = plan.assemble_km(1.0, 1.0, "fp64")
    sel = pec_solve_ids(plan, f, [], walls)
    sub = CscMap(plan).restrict(sel)
    data = sub.to_csc_combine(K, M, k0, block_symmetric=True)
    bvec = torch.zeros(f.n_field, dtype=torch.complex128, device="cuda")
    for tris, md, driven in ((p1, modes[0], True), (p2, modes[1], False)):
        rs = RobinSurface(plan, f, tris)
        rs.add_to_csc(data, rs.positions(sub, sel), rs.blocks(1.0), md.get_gamma(k0))
        if driven:
            x, y, z = generate_points_3d(f.mesh.nodes, f.mesh.tris, tris)
            U = md.get_Uinc(x, y, z, k0).reshape(3, DPTS.shape[1], len(tris))
            rs.force(U, bvec=bvec)
    ex = sub.export("indptr", "indices")
    A = to_scipy(ex["indptr"], ex["indices"], data, sel.n_solve)
    bsel = bvec[sel.solve_ids].cpu().numpy()
    xs = sla.spsolve(A, bsel)
    sol = scatter_solution(torch.as_tensor(xs, device="cuda"), sel).cpu().numpy()
    fieldf = interpolate_Ef(f, sol)
    tv1, tv2 = f.mesh.tris[:, p1], f.mesh.tris[:, p2]
    S11 = s_parameter(f.mesh.nodes, tv1, modes[0], True, k0, fieldf)
    S21 = s_parameter(f.mesh.nodes, tv2, modes[1], False, k0, fieldf)
    assert abs(beta - 158.238) < 2e-3
    assert abs(S11) < 0.05, S11
    assert abs(abs(S21) - 1) < 0.03, S21
    dphi = np.angle(S21 * np.exp(1j * beta * L_WG))
    assert abs(dphi) < 0.05, (np.angle(S21), -beta * L_WG)


def test_two_port_s_parameters_gpu_oracle_and_mixed(wr90_field):
    """The same two-port WR-90 problem (10 GHz) solved by SciPy's sparse direct solver on
    three volume systems: the GPU's FP64 assembly, the oracle's, and the GPU's mixed
    FP32/FP64 assembly (complex64 COO, FP64 CSC sums), each reduced to the free DOFs and
    completed with the same GPU port blocks. GPU and oracle give the same field and
    S-parameters (the end-to-end check of the volume assembly); the mixed path's errors
    against FP64 per matrix, per solved field and per S-parameter are printed and bounded."""
    import torch
    import oracle
    from emerge_amd.assembly import CscMap, Ned2Plan
    from emerge_amd.interp import interpolate_Ef
    from emerge_amd.pec import pec_solve_ids, scatter_solution, to_scipy
    from emerge_amd.ports import LocalFrame, RectangularWaveguideMode
    from emerge_amd.robin import DPTS, RobinSurface, generate_points_3d
    from emerge_amd.sparam import s_parameter
    from tests.conftest import ref_tensor
    f = wr90_field
    T = f.n_tets
    k0 = 2 * np.pi * 10e9 / C0
    plan = Ned2Plan.from_field(f)
    p1, p2 = _port_tris(f, 0.0), _port_tris(f, L_WG)
    walls = np.setdiff1d(f.mesh.boundary_tris(), np.concatenate([p1, p2]))
    modes = [RectangularWaveguideMode(LocalFrame([A_WG / 2, B_WG / 2, z], [1, 0, 0], [0, 1, 0],
                                                 [0, 0, 1]), A_WG, B_WG, (1, 0))
             for z in (0.0, L_WG)]
    sel = pec_solve_ids(plan, f, [], walls)
    sub = CscMap(plan).restrict(sel)
    ex = sub.export("indptr", "indices")
    ids = sel.solve_ids.cpu().numpy()
    # volume systems on the free DOFs
    K, M = plan.assemble_km(1.0, 1.0, "fp64")
    vol = sub.to_csc_combine(K, M, k0, block_symmetric=True)
    Km, Mm = plan.assemble_km(1.0, 1.0, "mixed")
    vol_m = sub.to_csc_combine(Km, Mm, k0, block_symmetric=True).to(torch.complex128)
    I = ref_tensor(T, "identity", None)
    dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(f, I, I, [])
    A_ref = oracle.CSCMap(rows, cols, f.n_field).to_csc(dE - dB * k0 ** 2)
    # port blocks and excitation (GPU), identical for the three systems
    full = vol.clone()
    bvec = torch.zeros(f.n_field, dtype=torch.complex128, device="cuda")
    for tris, md, driven in ((p1, modes[0], True), (p2, modes[1], False)):
        rs = RobinSurface(plan, f, tris)
        rs.add_to_csc(full, rs.positions(sub, sel), rs.blocks(1.0), md.get_gamma(k0))
        if driven:
            x, y, z = generate_points_3d(f.mesh.nodes, f.mesh.tris, tris)
            U = md.get_Uinc(x, y, z, k0).reshape(3, DPTS.shape[1], len(tris))
            rs.force(U, bvec=bvec)
    ports = to_scipy(ex["indptr"], ex["indices"], full - vol, sel.n_solve)
    bsel = bvec[sel.solve_ids].cpu().numpy()
    systems = {
        "gpu": to_scipy(ex["indptr"], ex["indices"], full, sel.n_solve),
        "oracle": (A_ref[ids][:, ids] + ports).tocsc(),
        "mixed": (to_scipy(ex["indptr"], ex["indices"], vol_m, sel.n_solve) + ports).tocsc(),
    }
    tv1, tv2 = f.mesh.tris[:, p1], f.mesh.tris[:, p2]
    sol, S = {}, {}
    for name, A in systems.items():
        xs = sla.spsolve(A, bsel)
        sol[name] = xs
        fld = interpolate_Ef(f, scatter_solution(torch.as_tensor(xs, device="cuda"),
                                                 sel).cpu().numpy())
        S[name] = np.array([s_parameter(f.mesh.nodes, tv1, modes[0], True, k0, fld),
                            s_parameter(f.mesh.nodes, tv2, modes[1], False, k0, fld)])
    assert abs(S["gpu"][0]) < 0.05 and abs(abs(S["gpu"][1]) - 1) < 0.03, S["gpu"]
    assert rel_fro(sol["gpu"], sol["oracle"]) < 1e-9
    assert np.abs(S["gpu"] - S["oracle"]).max() < 1e-9, (S["gpu"], S["oracle"])
    e_mat = rel_fro(vol_m.cpu().numpy(), vol.cpu().numpy())
    e_fld = rel_fro(sol["mixed"], sol["gpu"])
    e_s = np.abs(S["mixed"] - S["gpu"]).max()
    print(f"\n[two-port WR-90, 7,200 tets, 10 GHz] mixed vs FP64: free-DOF matrix {e_mat:.3e}, "
          f"solved field {e_fld:.3e}, max |dS| {e_s:.3e}; GPU vs oracle field "
          f"{rel_fro(sol['gpu'], sol['oracle']):.3e}, max |dS| "
          f"{np.abs(S['gpu'] - S['oracle']).max():.3e}")
    assert e_mat < 1e-6 and e_fld < 1e-3 and e_s < 1e-4
