"""The oracle's restatement of the plastic force response (VolumetricFemForceFpResponse: value f(Fp) = dE/dx at fixed
x and theta, Jacobian in the per-tet column-major vec(Fp)) pinned on the CPU: on the two-tet problem of the
reference's tests/python/test_fp_force_response.py the value equals the energy gradient of the oracle's independent
plastic energy path (atol 1e-13) and the Jacobian equals central differences of that gradient (rtol 1e-7, atol 1e-9,
the reference's own tolerances); the point derivative d2W_dF_dFp equals central differences of P = dW/dF in Fp; the
transpose product equals J^T w; and a dropped term or a flipped sign in d2W_dF_dFp fails the difference check."""
import numpy as np
import pytest


def problem():
    # tests/python/test_material_force_response.py:11-20 of the reference (theta of its Fp test: (2, 3), (4, 2)).
    nodes = np.array([[0., 0., 0.], [1.2, 0., 0.], [0., 1., 0.], [0., 0., .8], [1.2, 1., .8]])
    tets = np.array([[0, 1, 2, 3], [1, 2, 3, 4]], dtype=np.int32)
    F = np.array([[1.15, .12, 0.], [.03, .94, .08], [0., .02, 1.07]])
    u = (nodes @ (F - np.eye(3)).T).ravel()
    Fp = np.stack([np.array([[1.1 + .05 * e, 0., 0.], [0., 1., .1], [0., 0., 1.]]).ravel(order="F")
                   for e in range(2)])
    theta = np.array([[2., 3.], [4., 2.]])
    return nodes, tets, u, Fp, theta


def energy_gradient(oracle_mod, nodes, tets, theta, u, Fp):
    mats = np.asarray(Fp).reshape(-1, 3, 3).transpose(0, 2, 1)  # column-major vec -> matrices
    return oracle_mod.OracleEnergy(len(nodes), tets, nodes, theta, Fp=mats).eval(u, True)[1]


def test_value_is_the_energy_gradient(oracle_mod):
    nodes, tets, u, Fp, theta = problem()
    f, _, _ = oracle_mod.force_fp_products(tets, nodes, theta, u, Fp)
    np.testing.assert_allclose(f, energy_gradient(oracle_mod, nodes, tets, theta, u, Fp), atol=1e-13)


def test_jacobian_matches_central_differences_of_the_energy_gradient(oracle_mod):
    nodes, tets, u, Fp, theta = problem()
    n = Fp.size
    J = np.stack([oracle_mod.force_fp_products(tets, nodes, theta, u, Fp, dFp=np.eye(n)[c].reshape(Fp.shape))[2]
                  for c in range(n)], axis=1)
    h = 1e-6
    for c in range(n):
        d = h * np.eye(n)[c].reshape(Fp.shape)
        fd = (energy_gradient(oracle_mod, nodes, tets, theta, u, Fp + d) -
              energy_gradient(oracle_mod, nodes, tets, theta, u, Fp - d)) / (2 * h)
        np.testing.assert_allclose(J[:, c], fd, rtol=1e-7, atol=1e-9)
    w = np.random.default_rng(1).normal(size=u.size)
    _, jt, _ = oracle_mod.force_fp_products(tets, nodes, theta, u, Fp, w=w)
    np.testing.assert_allclose(jt.ravel(), J.T @ w, rtol=1e-13, atol=1e-14)


def point_fd(oracle_mod, F, Fp, theta, h=1e-6):
    out = np.zeros((9, 9))
    for c in range(9):
        d = np.zeros(9)
        d[c] = h
        d = d.reshape(3, 3, order="F")
        Pp = oracle_mod.neohookean(F, Fp + d, theta, A=False, mixed=False)[1]
        Pm = oracle_mod.neohookean(F, Fp - d, theta, A=False, mixed=False)[1]
        out[:, c] = ((Pp - Pm) / (2 * h)).reshape(-1, order="F")
    return out


def test_point_derivative_matches_central_differences(oracle_mod):
    rng = np.random.default_rng(7)
    for _ in range(20):
        F = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
        Fp = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
        if np.linalg.det(F) <= 0.1 or np.linalg.det(Fp) <= 0.1:
            continue
        theta = np.array([rng.uniform(0.5, 5), rng.uniform(0.5, 5)])
        M = oracle_mod.d2W_dF_dFp(F, Fp, theta)
        np.testing.assert_allclose(M, point_fd(oracle_mod, F, Fp, theta), rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("mistake", ["drop_volume_term", "flip_last_sign"])
def test_planted_mistake_fails_the_difference_check(oracle_mod, mistake):
    F = np.array([[1.15, .12, 0.], [.03, .94, .08], [0., .02, 1.07]])
    Fp = np.array([[1.1, 0., 0.], [0., 1., .1], [0., 0., 1.]])
    theta = np.array([2., 3.])
    good = oracle_mod.d2W_dF_dFp(F, Fp, theta)
    # The planted variants, built from the same pieces with one term changed.
    inv_Fp = np.linalg.inv(Fp)
    Fe = F @ inv_Fp
    P = oracle_mod.neohookean(F, Fp, theta, A=False, mixed=False)[1]
    Ae = oracle_mod.neohookean(Fe, np.eye(3), theta, P=False, mixed=False)[2]
    bad = np.zeros((9, 9))
    for c in range(9):
        dFp = np.eye(9)[c].reshape(3, 3, order="F")
        dS = (Ae @ (-Fe @ dFp @ inv_Fp).reshape(-1, order="F")).reshape(3, 3, order="F")
        vol = 0.0 if mistake == "drop_volume_term" else np.trace(inv_Fp @ dFp)
        sign = 1.0 if mistake == "flip_last_sign" else -1.0
        dP = vol * P + np.linalg.det(Fp) * dS @ inv_Fp.T + sign * P @ dFp.T @ inv_Fp.T
        bad[:, c] = dP.reshape(-1, order="F")
    fd = point_fd(oracle_mod, F, Fp, theta)
    np.testing.assert_allclose(good, fd, rtol=1e-6, atol=1e-7)
    assert not np.allclose(bad, fd, rtol=1e-6, atol=1e-7)


def test_reference_known_answers(oracle_mod):
    """tests/cpp/fem/material/volumetric/neo_hookean_gtest.cpp:270-300 of the reference: at F = Fp = I the mixed
    derivative is minus the tangent; at F = a I, Fp = b I it maps vec(I) to (mu a + b^2/a (9 lambda log(a/b)
    - 3 mu - 3 lambda)) vec(I); at its fixed (F, Fp) it matches central differences of P (1e-8); and a zero Fp or an
    inverted F is a domain error, a non-finite Fp or mu = 0 an invalid argument."""
    I = np.eye(3)
    theta = np.array([2.0, 3.0])
    A = oracle_mod.neohookean(I, I, theta, P=False, mixed=False)[2]
    np.testing.assert_allclose(oracle_mod.d2W_dF_dFp(I, I, theta), -A, rtol=0, atol=1e-14 * np.abs(A).max())
    a, b = 1.3, 0.8
    B = oracle_mod.d2W_dF_dFp(a * I, b * I, theta)
    expected = theta[1] * a + b * b / a * (9 * theta[0] * np.log(a / b) - 3 * theta[1] - 3 * theta[0])
    np.testing.assert_allclose(B @ I.ravel(order="F"), expected * I.ravel(order="F"), rtol=1e-13, atol=1e-13)
    F = np.array([[1.2, .2, -.1], [.05, .9, .1], [.02, -.04, 1.1]])
    Fp = np.array([[.9, -.1, .03], [.04, 1.2, .08], [-.02, .05, 1.1]])
    M = oracle_mod.d2W_dF_dFp(F, Fp, theta)
    fd = point_fd(oracle_mod, F, Fp, theta)
    for c in range(9):
        assert np.linalg.norm(M[:, c] - fd[:, c]) <= 1e-8 * max(np.linalg.norm(fd[:, c]), np.linalg.norm(M[:, c]))
    nan_Fp = Fp.copy()
    nan_Fp[0, 0] = np.nan
    for args, code in (((F, 1e-300 * I, theta), 2), ((-F, Fp, theta), 2), ((F, Fp, np.array([2.0, 0.0])), 1),
                       ((F, nan_Fp, theta), 1)):
        with pytest.raises(oracle_mod.OracleError) as err:
            oracle_mod.d2W_dF_dFp(*args)
        assert err.value.code == code
