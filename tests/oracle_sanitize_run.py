"""Calls every C entry point of the oracle on two small meshes and saves the outputs.

Run as a child process by tests/test_oracle_sanitize.py: once with the normal build and
once with the AddressSanitizer/UndefinedBehaviorSanitizer build (EMERGE_ORACLE_LIB, the
ASan runtime preloaded). Usage: ``python tests/oracle_sanitize_run.py OUT.npz``.
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import oracle  # noqa: E402
from oracle import heat as oh  # noqa: E402
from oracle import interp as oi  # noqa: E402
from oracle import robin as orob  # noqa: E402
from emerge_amd.mesh import (Legrange2Field, Nedelec2Field, delaunay_box_mesh,  # noqa: E402
                             hex_box_mesh)


def _tensor(T, rng):
    a = rng.normal(size=(3, 3, T)) + 1j * rng.normal(size=(3, 3, T))
    return a + 4 * np.eye(3)[:, :, None]


def _interior_points(mesh, rng, n):
    t = rng.integers(0, mesh.n_tets, n)
    w = rng.dirichlet(np.ones(4), n)
    pts = np.einsum("nk,dkn->dn", w, mesh.nodes[:, mesh.tets[:, t]])
    return pts


def main(out_path):
    rng = np.random.default_rng(5)
    res = {"keast": oracle.keast_table(), "dpts": orob.dpts()}
    meshes = {"hex": hex_box_mesh(2e-2, 1e-2, 3e-2, 3, 2, 4),
              "del": delaunay_box_mesh(300, seed=3)}
    for name, m in meshes.items():
        f = Nedelec2Field(m)
        T = f.n_tets
        er, ur = _tensor(T, rng), _tensor(T, rng)
        cond = np.array([1, T // 2], np.int64)
        for lf in (False, True):
            dE, dB, rows, cols = oracle.tet_mass_stiffness_matrices(
                f, er, ur, cond, nthreads=2, local_frame=lf)
            res[f"{name}_E_{int(lf)}"], res[f"{name}_B_{int(lf)}"] = dE, dB
        res[f"{name}_rows"], res[f"{name}_cols"] = rows, cols
        cm = oracle.CSCMap(rows, cols, f.n_field)
        res[f"{name}_indptr"], res[f"{name}_indices"] = cm.indptr, cm.indices
        res[f"{name}_csc"] = cm.to_csc_data(dE - 40.0 * dB)
        pts = _interior_points(m, rng, 64)
        tet_of = oi.locate(m, pts)
        sol = rng.normal(size=f.n_field) + 1j * rng.normal(size=f.n_field)
        res[f"{name}_loc"] = tet_of
        res[f"{name}_E"] = oi.interp(f, pts, sol, tet_of)
        res[f"{name}_curlE"] = oi.interp(f, pts, sol, tet_of, curl=True, local_frame=True)
        surf = m.boundary_tris()[:40]
        res[f"{name}_bc"] = orob.bc_entries(m.nodes, m.tris, surf, 0.3 - 1.2j)
        U = rng.normal(size=(3, 6, surf.size)) + 1j * rng.normal(size=(3, 6, surf.size))
        C = rng.normal(size=(3, 6, surf.size)) + 1j * rng.normal(size=(3, 6, surf.size))
        nrm = rng.normal(size=(3, surf.size))
        res[f"{name}_force"] = orob.force_entries(m.nodes, m.tris, f.tri_to_field, f.n_field,
                                                  surf, U)
        res[f"{name}_force_scat"] = orob.force_entries(m.nodes, m.tris, f.tri_to_field,
                                                       f.n_field, surf, U, U_curl=C,
                                                       normals=nrm)
        h = Legrange2Field(m)
        hc = np.abs(rng.normal(size=(3, 3, T))) + 2 * np.eye(3)[:, :, None]
        K, M, hr, hcol = oh.build(m.nodes, m.tets, h.tet_to_field, cond=hc,
                                  rho_cp=rng.uniform(1, 2, T), rowcol=True)
        res[f"{name}_heatK"], res[f"{name}_heatM"] = K, M
        res[f"{name}_heat_rows"], res[f"{name}_heat_cols"] = hr, hcol
    np.savez(out_path, **res)
    print(f"oracle library: {oracle._LIB_PATH}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1])
