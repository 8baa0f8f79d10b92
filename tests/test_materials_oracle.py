"""Pins of the material-preparation oracle (oracle/materials.py, restating
assembler.py:439-469) by the textbook complex permittivity, special cases and a
hand-written table of the comparison rules."""
import numpy as np

from oracle import materials as om

F = 10e9
D = 2 * np.pi * F * om.EPS0  # omega * eps0


def _rand(rng, T, scale=1.0):
    return scale * (rng.normal(size=(3, 3, T)) + 1j * rng.normal(size=(3, 3, T)))


def test_textbook_complex_permittivity():
    """Real isotropic eps', tan d, sigma: eps = eps'(1 - j tan d) - j sigma/(w eps0)."""
    rng = np.random.default_rng(0)
    T = 500
    ep, td, sg = rng.uniform(1, 10, T), rng.uniform(0, 0.02, T), rng.uniform(0, 50, T)
    er = np.zeros((3, 3, T), complex)
    tand = np.zeros_like(er)
    cond = np.zeros_like(er)
    for d in range(3):
        er[d, d], tand[d, d], cond[d, d] = ep, td, sg
    out, ids, fd = om.prepare_materials(er, tand, cond, F)
    want = ep - 1j * (ep * td + sg / D)
    for d in range(3):
        assert np.array_equal(out[d, d].real, ep)
        assert np.all(np.abs(out[d, d].imag - want.imag) <= 4e-16 * np.abs(want.imag))
    assert np.all(out[0, 1] == 0) and np.all(out[2, 0] == 0)
    assert ids.size == 0 and fd  # 0 < sigma < 1e8 somewhere


def test_lossless_unchanged_and_conductivity_alone():
    rng = np.random.default_rng(1)
    T = 300
    er = _rand(rng, T, 3.0)
    z = np.zeros_like(er)
    out, ids, fd = om.prepare_materials(er, z, z, F)
    assert np.array_equal(out, er) and ids.size == 0 and not fd
    # er = tand = 0: -j (a + j b) / d = (b - j a) / d
    cond = _rand(rng, T, 100.0)
    out, _, _ = om.prepare_materials(z, z, cond, F)
    want = (cond.imag - 1j * cond.real) / D
    assert np.all(np.abs(out - want) <= 4e-16 * np.abs(want))


def test_anisotropic_complex_loss_tangent():
    """Full complex tensors against Python's complex arithmetic (no FMA; the two differ in
    the last bit at most)."""
    rng = np.random.default_rng(2)
    T = 64
    er, tand, cond = _rand(rng, T, 4.0), _rand(rng, T, 0.05), _rand(rng, T, 10.0)
    out, _, _ = om.prepare_materials(er, tand, cond, F)
    for idx in np.ndindex(3, 3, T):
        w = complex(er[idx]) * (1 - 1j * complex(tand[idx])) - 1j * complex(cond[idx]) / D
        assert abs(complex(out[idx]) - w) <= 1e-15 * max(abs(w), 1e-300), idx


def test_conductor_and_frequency_rules():
    """cond[0,0] of each tet against the two limits, complex values ordered real part first;
    a large conductivity in another component does not make a conductor."""
    vals = [(0.0, False, False), (5.0, False, True), (1e4, False, True),
            (1e4 + 1j, True, True), (1e4 - 1j, False, True), (10000.001, True, True),
            (1e8, True, False), (1e8 + 1j, True, False), (1e8 - 1j, True, True),
            (2e9, True, False), (1j, False, True), (-1j, False, False), (-3.0, False, False)]
    T = len(vals)
    cond = np.zeros((3, 3, T), complex)
    cond[0, 0] = [v for v, _, _ in vals]
    z = np.zeros_like(cond)
    _, ids, _ = om.prepare_materials(z, z, cond, F)
    assert ids.tolist() == [t for t, (_, c, _) in enumerate(vals) if c]
    z1 = np.zeros((3, 3, 1), complex)
    for v, _, f in vals:
        one = z1.copy()
        one[0, 0, 0] = v
        assert om.prepare_materials(z1, z1, one, F)[2] == f, v
    big = np.zeros((3, 3, 2), complex)
    big[1, 1, 0] = 1e9
    assert om.prepare_materials(z[:, :, :2], z[:, :, :2], big, F)[1].size == 0
