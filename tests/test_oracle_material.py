"""The oracle's NeoHookean material against the closed forms and identities of the reference's own tests
(tests/cpp/fem/material/volumetric/neo_hookean_gtest.cpp), plus central finite differences."""
import math

import numpy as np
import pytest


def rotation(angle, axis):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


I = np.eye(3)
FP = np.array([[1.3, 0.2, 0.1], [0.0, 0.9, -0.1], [0.1, 0.2, 1.2]])


def test_stress_free_and_linear_elastic_tangent(oracle_mod):
    # neo_hookean_gtest.cpp:93-110
    theta = [2.0, 3.0]
    R = rotation(0.7, [1, 2, 3])
    for F in (FP, R @ FP):
        W, P, _, _ = oracle_mod.neohookean(F, FP, theta)
        assert abs(W) < 1e-13
        assert np.linalg.norm(P) < 1e-13
    _, _, A, _ = oracle_mod.neohookean(I, I, theta)
    for j in range(9):
        D = np.zeros(9)
        D[j] = 1
        D = D.reshape(3, 3, order="F")
        expected = theta[1] * (D + D.T) + theta[0] * np.trace(D) * I
        assert np.linalg.norm(A[:, j] - expected.reshape(-1, order="F")) < 1e-13


def test_analytic_dilation_and_objectivity(oracle_mod):
    # neo_hookean_gtest.cpp:112-126
    theta = [2.0, 3.0]
    s, p = 1.25, 1.3
    L = 3 * math.log(s)
    psi = 1.5 * theta[1] * (s * s - 1) - theta[1] * L + 0.5 * theta[0] * L * L
    W, P, _, _ = oracle_mod.neohookean(s * p * I, p * I, theta)
    assert abs(W - p ** 3 * psi) < 1e-13
    expected_P = p * p * (theta[1] * s + (theta[0] * L - theta[1]) / s) * I
    assert np.allclose(P, expected_P, rtol=1e-13, atol=0)
    F = np.array([[1.2, 0.2, 0.1], [-0.1, 0.8, 0.3], [0.2, 0.1, 1.1]])
    R = rotation(0.6, [2, -1, 3])
    W1, P1, _, _ = oracle_mod.neohookean(R @ F, p * I, theta)
    W0, P0, _, _ = oracle_mod.neohookean(F, p * I, theta)
    assert abs(W1 - W0) < 1e-13
    assert np.linalg.norm(P1 - R @ P0) <= 1e-13 * np.linalg.norm(P1)


def test_mixed_parameter_derivative_closed_forms(oracle_mod):
    # neo_hookean_gtest.cpp:15-33
    theta = [0.0, 3.0]
    p, s = 1.3, 1.25
    _, _, _, M = oracle_mod.neohookean(p * s * I, p * I, theta)
    expected = np.zeros((9, 2))
    for i in range(3):
        expected[4 * i, 0] = p * p * 3 * math.log(s) / s
        expected[4 * i, 1] = p * p * (s - 1 / s)
    assert np.abs(M - expected).max() <= 1e-13 * np.abs(expected).max()
    F = I.copy()
    F[0, 1] = 0.4
    _, _, _, M = oracle_mod.neohookean(F, I, theta)
    expected = np.zeros((9, 2))
    expected[3, 1] = 0.4
    expected[1, 1] = 0.4
    assert np.abs(M - expected).max() < 1e-14
    _, _, _, M = oracle_mod.neohookean(p * I, p * I, theta)
    assert np.linalg.norm(M) < 1e-13


def test_mixed_derivative_reconstructs_stress(oracle_mod):
    # neo_hookean_gtest.cpp:35-52: B theta == vec P exactly (P is linear in theta), and R-equivariance.
    F = np.array([[1.2, 0.3, -0.1], [0.1, 0.9, 0.2], [-0.05, 0.1, 1.1]])
    _, _, _, B = oracle_mod.neohookean(F, FP, [2.0, 3.0])
    for theta in ([0.0, 1.0], [5.0, 0.7], [2.0, 3.0]):
        _, P, _, B2 = oracle_mod.neohookean(F, FP, theta)
        assert np.allclose(B @ np.array(theta), P.reshape(-1, order="F"), rtol=1e-13, atol=1e-15)
        assert np.array_equal(B, B2)
    R = rotation(0.7, [1, 2, 3])
    _, _, _, Br = oracle_mod.neohookean(R @ F, FP, [2.0, 3.0])
    for c in range(2):
        expected = R @ B[:, c].reshape(3, 3, order="F")
        assert np.allclose(Br[:, c], expected.reshape(-1, order="F"), rtol=1e-13, atol=1e-15)


def test_full_gradient_and_hessian_finite_differences(oracle_mod):
    # neo_hookean_gtest.cpp:128-179 (12 varied states, compression and det(Fp) != 1)
    theta = [2.0, 3.0]
    for sample in range(12):
        Fp = np.zeros((3, 3))
        Fe = np.zeros((3, 3))
        for j in range(9):
            Fp.reshape(-1, order="F")  # noqa: B018 (documentation of the column-major fill below)
        fp = np.array([0.12 * math.sin(1.7 * (j + 1) * (sample + 1)) for j in range(9)]).reshape(3, 3, order="F")
        fe = np.array([0.15 * math.cos(0.8 * (j + 1) * (sample + 2)) for j in range(9)]).reshape(3, 3, order="F")
        Fp = fp + (0.8 + 0.06 * sample) * I
        Fe = fe + (0.7 + 0.05 * sample) * I
        F = Fe @ Fp
        W, P, A, _ = oracle_mod.neohookean(F, Fp, theta)
        h = 1e-6
        fdP = np.zeros(9)
        fdA = np.zeros((9, 9))
        for k in range(9):
            d = np.zeros(9)
            d[k] = h
            D = d.reshape(3, 3, order="F")
            Wp, Pp, _, _ = oracle_mod.neohookean(F + D, Fp, theta)
            Wm, Pm, _, _ = oracle_mod.neohookean(F - D, Fp, theta)
            fdP[k] = (Wp - Wm) / (2 * h)
            fdA[:, k] = (Pp - Pm).reshape(-1, order="F") / (2 * h)
        assert np.linalg.norm(fdP - P.reshape(-1, order="F")) <= 1e-7 * (1 + np.linalg.norm(P))
        assert np.linalg.norm(fdA - A) <= 1e-6 * (1 + np.linalg.norm(A))
        assert np.abs(A - A.T).max() == 0.0 or np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()


def test_error_classes(oracle_mod):
    # neo_hookean_gtest.cpp:54-91 and neo_hookean.cpp:14-31
    th = [2.0, 3.0]
    for bad in ([-1.0, 3.0], [2.0, 0.0], [2.0, math.inf]):
        with pytest.raises(oracle_mod.OracleError) as e:
            oracle_mod.neohookean(I, I, bad)
        assert e.value.code == 1
    Fb = I.copy()
    Fb[0, 0] = math.nan
    for F, Fp in ((Fb, I), (I, Fb)):
        with pytest.raises(oracle_mod.OracleError) as e:
            oracle_mod.neohookean(F, Fp, th)
        assert e.value.code == 1
    for F, Fp in ((-I, I), (I, np.zeros((3, 3)))):
        with pytest.raises(oracle_mod.OracleError) as e:
            oracle_mod.neohookean(F, Fp, th)
        assert e.value.code == 2
    F = I.copy()
    F[0, 0] = 1e155
    with pytest.raises(oracle_mod.OracleError) as e:
        oracle_mod.neohookean(F, I, th, P=False, A=False, mixed=False)
    assert e.value.code == 2
    F = I.copy()
    F[0, 0] = 2
    W, _, _, _ = oracle_mod.neohookean(F, I, [0.0, 1e308], P=False, A=False, mixed=False)
    assert math.isfinite(W)
    with pytest.raises(oracle_mod.OracleError) as e:
        oracle_mod.neohookean(F, I, [0.0, 1e308], P=True, A=False, mixed=False)
    assert e.value.code == 2
    W, P, _, _ = oracle_mod.neohookean(I, I, [0.0, 1e308], P=True, A=False, mixed=False)
    assert W == 0 and np.all(np.isfinite(P))
    with pytest.raises(oracle_mod.OracleError) as e:
        oracle_mod.neohookean(I, I, [0.0, 1e308], P=False, A=True, mixed=False)
    assert e.value.code == 2


def test_projection_known_answers(oracle_mod):
    # mesh_energy_gtest.cpp:142-155 (a local -2 projects to 0, a local 3 stays), :254-264 (overflow is a
    # domain error), and the clamp against LAPACK on random symmetric matrices.
    assert oracle_mod.project_psd(np.array([[-2.0]]))[0, 0] == 0.0
    assert oracle_mod.project_psd(np.array([[3.0]]))[0, 0] == 3.0
    with pytest.raises(oracle_mod.OracleError) as e:
        oracle_mod.project_psd(np.full((2, 2), np.finfo(float).max))
    assert e.value.code == 2
    rng = np.random.default_rng(3)
    for n in (2, 5, 9, 12):
        for _ in range(5):
            M = rng.normal(size=(n, n))
            lam, V = np.linalg.eigh(0.5 * (M + M.T))
            expected = (V * np.maximum(lam, 0)) @ V.T
            got = oracle_mod.project_psd(M)
            assert np.abs(got - expected).max() <= 1e-13 * np.abs(expected).max()
    # A PSD matrix is a fixed point up to rounding.
    M = rng.normal(size=(12, 12))
    S = M @ M.T
    assert np.abs(oracle_mod.project_psd(S) - S).max() <= 1e-13 * np.abs(S).max()
