"""CPU pins of the heat-conduction (Lagrange-2) oracle."""
import numpy as np
import pytest

from oracle import heat as oh


def _tet(rng, positive=True):
    v = rng.normal(size=(3, 4))
    J = np.array([v[:, k] - v[:, 0] for k in (1, 2, 3)]).T
    if (np.linalg.det(J) > 0) != positive:
        v[:, [2, 3]] = v[:, [3, 2]]
    return v


@pytest.mark.parametrize("seed", range(4))
def test_exact_rule_gives_exact_integrals(seed):
    """With the closed-form Keast rule the restatement equals exact integration (degree
    <= 4); with the reference's 8-digit table it differs by the table's rounding only."""
    rng = np.random.default_rng(seed)
    v = _tet(rng)
    C = rng.normal(size=(3, 3)) + 3 * np.eye(3)
    t, f = np.arange(4).reshape(4, 1), np.arange(10).reshape(10, 1)
    Ke, Me = oh.exact_element(v, C, 2.5)
    K, M, _, _ = oh.build(v, t, f, cond=C.reshape(3, 3, 1), rho_cp=np.array([2.5]),
                          rule=oh.exact_rule())
    assert np.abs(K.reshape(10, 10) - Ke).max() <= 1e-13 * np.abs(Ke).max()
    assert np.abs(M.reshape(10, 10) - Me).max() <= 1e-13 * np.abs(Me).max()
    K, M, _, _ = oh.build(v, t, f, cond=C.reshape(3, 3, 1), rho_cp=np.array([2.5]))
    dK = np.abs(K.reshape(10, 10) - Ke).max() / np.abs(Ke).max()
    dM = np.abs(M.reshape(10, 10) - Me).max() / np.abs(Me).max()
    assert 1e-9 < dK < 1e-6 and 1e-9 < dM < 1e-6


def test_rule_tables():
    R = oh.REFERENCE_RULE
    assert abs(R[0].sum() - (1 + 3e-7)) < 1e-12  # the 8-digit weights
    assert np.allclose(R[1:].sum(axis=0), 1.0, rtol=0, atol=1e-15)
    E = oh.exact_rule()
    assert abs(E[0].sum() - 1) < 1e-15
    assert np.abs(E - R).max() < 5e-7  # the centroid weight is rounded to 6 digits


def test_stiffness_rows_sum_to_zero_and_mass_total():
    """K annihilates constants for any rule; sum M = rho V sum_w (the rule's weights)."""
    rng = np.random.default_rng(9)
    v = _tet(rng)
    C = np.diag([1.0, 2.0, 3.0])
    t, f = np.arange(4).reshape(4, 1), np.arange(10).reshape(10, 1)
    K, M, _, _ = oh.build(v, t, f, cond=C.reshape(3, 3, 1), rho_cp=np.array([1.5]))
    K = K.reshape(10, 10)
    assert np.abs(K.sum(axis=1)).max() <= 1e-13 * np.abs(K).max()
    V = abs(np.linalg.det(np.array([v[:, k] - v[:, 0] for k in (1, 2, 3)]).T)) / 6
    assert abs(M.sum() - 1.5 * V * oh.REFERENCE_RULE[0].sum()) <= 1e-14 * 1.5 * V


def test_negative_orientation_follows_the_reference():
    """The reference divides by |V|: on a negatively oriented tet its functions are the
    Lagrange ones in l = -L. The oracle (restating that) gives the exact matrices of those
    functions, which differ from the true Lagrange ones."""
    rng = np.random.default_rng(10)
    v = _tet(rng, positive=False)
    t, f = np.arange(4).reshape(4, 1), np.arange(10).reshape(10, 1)
    _, M, _, _ = oh.build(v, t, f, rho_cp=np.array([1.0]), rule=oh.exact_rule())
    _, Me = oh.exact_element(v, np.eye(3), 1.0)
    assert np.abs(M.reshape(10, 10) - Me).max() > 1e-3 * np.abs(Me).max()
