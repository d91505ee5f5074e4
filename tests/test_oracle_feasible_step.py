"""The oracle's feasible step against the reference's known answers and invariants:
positive_det_feasible_step_gtest.cpp (every case below, same tolerances) and the Tet P1 energy's
FeasibleStepBoundsTheWholeSegmentAndUsesBothElements (volumetric_fem_energy_tet_p1_gtest.cpp:132-151), plus the
safety property that every point of [0, alpha] keeps det > 0."""
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"
I = np.eye(3)


def check_segment(A, dA, alpha):
    assert 0 <= alpha <= 1
    for i in range(101):
        d = np.linalg.det(A + (alpha * i / 100) * dA)
        assert d > 0 and np.isfinite(d)


def test_full_step_directions(oracle_mod):
    shear = np.zeros((3, 3))
    shear[0, 1] = 10
    rot = np.zeros((3, 3))
    rot[0, 1], rot[1, 0] = -4, 4
    for dA in (np.zeros((3, 3)), I, shear, rot):
        assert oracle_mod.det_step(I, dA) == 1
        check_segment(I, dA, 1)


def test_linear_quadratic_cubic(oracle_mod):
    for diag in ([-4, 0, 0], [-4, -2, 0], [-4, -2, -1]):
        dA = np.diag(diag).astype(float)
        alpha = oracle_mod.det_step(I, dA)
        assert abs(alpha - 0.225) <= 1e-12
        check_segment(I, dA, alpha)
    assert oracle_mod.det_step(I, -0.5 * I) == 1
    cyclic = np.array([[0, 1, 0], [0, 0, 1], [-8, 0, 0]], dtype=float)  # det(I + t C) = 1 - 8 t^3
    alpha = oracle_mod.det_step(I, cyclic)
    assert abs(alpha - 0.45) <= 1e-12
    check_segment(I, cyclic, alpha)


def test_interior_double_and_triple_roots(oracle_mod):
    for diag in ([-4, -4, 0], [-4, -4, 1], [-4, -4, -4]):
        dA = np.diag(diag).astype(float)
        assert np.linalg.det(I + 0.25 * dA) == 0
        alpha = oracle_mod.det_step(I, dA)
        assert abs(alpha - 0.225) <= 2e-5 and alpha < 0.25
        check_segment(I, dA, alpha)


def test_scaling_and_change_of_basis(oracle_mod):
    axis = np.array([1.0, 2.0, 3.0]) / np.sqrt(14.0)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    R = I + np.sin(0.7) * K + (1 - np.cos(0.7)) * K @ K  # Rodrigues rotation by 0.7 about (1, 2, 3)
    B = np.array([[1.2, 0.2, -0.1], [0.1, 0.9, 0.2], [0.02, -0.1, 1.1]])
    D = np.diag([-4.0, -2.0, 1.0])
    for scale in (1e-50, 0.1, 1.0, 10.0, 1e50):
        A, dA = scale * R @ B, scale * R @ D @ B
        alpha = oracle_mod.det_step(A, dA)
        assert abs(alpha - 0.225) <= 1e-12
        check_segment(A, dA, alpha)
    extreme = np.zeros((3, 3))
    extreme[0, 0] = -1e200
    tiny = oracle_mod.det_step(I, extreme)
    assert abs(tiny / 9e-201 - 1) <= 1e-12
    assert oracle_mod.det_step(I, 1e300 * I) == 0


def test_rejects_invalid_starts_and_directions(oracle_mod):
    Z = np.zeros((3, 3))
    for A, dA, code in ((Z, I, 2), (-I, Z, 2), (1e200 * I, Z, 2)):
        with pytest.raises(oracle_mod.OracleError) as e:
            oracle_mod.det_step(A, dA)
        assert e.value.code == code
    for value in (np.inf, np.nan):
        bad = I.copy()
        bad[0, 0] = value
        for A, dA in ((bad, Z), (I, bad)):
            with pytest.raises(oracle_mod.OracleError) as e:
                oracle_mod.det_step(A, dA)
            assert e.value.code == 1


def test_two_tet_energy_feasible_step(oracle_mod):
    d = np.load(GOLD / "two_tet.npz")
    X = d["X"].reshape(-1)
    E = oracle_mod.OracleEnergy(5, d["tets"], d["X"], d["theta"])
    zero = np.zeros(15)
    assert E.max_feasible_step(zero, zero) == 1
    assert E.max_feasible_step(zero, 0.1 * X) == 1
    direction = zero.copy()
    direction[0::3] = -4 * X[0::3]
    direction[1::3] = -4 * X[1::3]
    alpha = E.max_feasible_step(zero, direction)
    assert abs(alpha - 0.225) <= 1e-6
    for i in range(21):
        assert np.isfinite(E.eval((alpha * i / 20) * direction)[0])
    direction[:] = 0
    direction[14] = 16
    assert abs(E.max_feasible_step(zero, direction) - 0.1125) <= 1e-12
    with pytest.raises(oracle_mod.OracleError) as e:
        E.max_feasible_step(-X, zero)
    assert e.value.code == 2
    with pytest.raises(oracle_mod.OracleError) as e:
        E.max_feasible_step(zero, np.full(15, np.nan))
    assert e.value.code == 1
