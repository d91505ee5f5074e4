"""The oracle's assembled energy against the reference's Tet P1 tests
(tests/cpp/fem/assembly/volumetric_fem_energy_tet_p1_gtest.cpp), brute-force patterns and determinism."""
import itertools
import math

import numpy as np
import pytest

from litepgo_amd.fem import kuhn_grid, lame_parameters

X2 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0.3, 0.2, -2]], dtype=np.float64)
T2 = np.array([[0, 1, 2, 3], [0, 2, 1, 4]], dtype=np.int32)
TH2 = np.array([[2.0, 3.0]] * 2)


def states():
    for state in range(3):
        x = 0.1 * X2.reshape(-1)
        if state == 0:
            x = np.zeros(15)
        if state == 2:
            x = x + 0.02 * np.sin(np.arange(15) + 0.3)
        yield x


def dense(E, H):
    outer, inner = E.pattern()
    import scipy.sparse as sp

    return sp.csc_matrix((H, inner, outer), shape=(E.num_dofs, E.num_dofs)).toarray()


def test_two_tet_energy_closed_form_and_finite_differences(oracle_mod):
    E = oracle_mod.OracleEnergy(5, T2, X2, TH2)
    assert np.allclose(E.weights(), [1 / 6, 1 / 3], rtol=1e-15)
    for x in states():
        value, g, Hv = E.eval(x, True, "exact")
        H = dense(E, Hv)
        assert np.linalg.norm(H - H.T) < 1e-12 * (1 + np.linalg.norm(H))
        expected = 0.0
        for cell in T2:
            Dm = np.stack([X2[cell[c + 1]] - X2[cell[0]] for c in range(3)], 1)
            xx = x.reshape(5, 3)
            Du = np.stack([xx[cell[c + 1]] - xx[cell[0]] for c in range(3)], 1)
            W, _, _, _ = oracle_mod.neohookean(np.eye(3) + Du @ np.linalg.inv(Dm), np.eye(3), [2, 3], P=False,
                                               A=False, mixed=False)
            expected += np.linalg.det(Dm) / 6 * W
        assert abs(value - expected) < 1e-13
        for h in (1e-3, 3e-4, 1e-4):
            fd_g = np.zeros(15)
            fd_H = np.zeros((15, 15))
            for j in range(15):
                p, m = x.copy(), x.copy()
                p[j] += h
                m[j] -= h
                vp, gp, _ = E.eval(p, True)
                vm, gm, _ = E.eval(m, True)
                fd_g[j] = (vp - vm) / (2 * h)
                fd_H[:, j] = (gp - gm) / (2 * h)
            assert np.linalg.norm(g - fd_g) / (1 + np.linalg.norm(g)) < 10 * h * h + 1e-9
            assert np.linalg.norm(H - fd_H) / (1 + np.linalg.norm(H)) < 10 * h * h + 1e-9


def test_invalid_geometry_is_rejected(oracle_mod):
    # volumetric_fem_energy_tet_p1_gtest.cpp:117-130
    for scale in (0.0, -1.0, 1e200):
        with pytest.raises(oracle_mod.OracleError) as e:
            oracle_mod.OracleEnergy(5, T2, scale * X2, TH2)
        assert e.value.code == 1
    for bad in (math.nan, math.inf, -math.inf):
        X = X2.copy()
        X[0, 0] = bad
        with pytest.raises(oracle_mod.OracleError) as e:
            oracle_mod.OracleEnergy(5, T2, X, TH2)
        assert e.value.code == 1
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.OracleEnergy(5, np.array([[0, 1, 2, 2]]), X2, TH2[:1])
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.OracleEnergy(5, np.array([[0, 1, 2, 5]]), X2, TH2[:1])


def test_eval_error_classes(oracle_mod):
    E = oracle_mod.OracleEnergy(5, T2, X2, TH2)
    x = np.zeros(15)
    x[0] = math.nan
    with pytest.raises(oracle_mod.OracleError) as e:
        E.eval(x)
    assert e.value.code == 1
    inverted = -2.0 * X2.reshape(-1)  # F = -I on both tets
    with pytest.raises(oracle_mod.OracleError) as e:
        E.eval(inverted, True, "exact")
    assert e.value.code == 2
    bad = oracle_mod.OracleEnergy(5, T2, X2, np.array([[2.0, 3.0], [-1.0, 3.0]]))
    with pytest.raises(oracle_mod.OracleError) as e:
        bad.eval(np.zeros(15))
    assert e.value.code == 1
    # the object stays usable after an error
    assert E.eval(np.zeros(15))[0] == 0.0


def test_union_of_elements_is_the_sum_of_single_element_energies(oracle_mod):
    # volumetric_fem_energy_tet_p1_gtest.cpp:196-248
    X, tets = kuhn_grid(2)
    lam, mu = lame_parameters(1e5, 0.4)
    theta = np.tile([[float(lam), float(mu)]], (len(tets), 1))
    x = np.random.default_rng(1).uniform(-0.02, 0.02, X.size)
    E = oracle_mod.OracleEnergy(len(X), tets, X, theta)
    value, g, Hv = E.eval(x, True, "exact")
    H = dense(E, Hv)
    total, gsum, Hsum = 0.0, np.zeros_like(g), np.zeros_like(H)
    for e in range(len(tets)):
        Ee = oracle_mod.OracleEnergy(len(X), tets[e:e + 1], X, theta[e:e + 1])
        v1, g1, H1 = Ee.eval(x, True, "exact")
        total += v1
        gsum += g1
        Hsum += dense(Ee, H1)
    assert abs(value - total) <= 1e-13 * abs(total)
    assert np.abs(g - gsum).max() <= 1e-13 * np.abs(g).max()
    assert np.abs(H - Hsum).max() <= 1e-13 * np.abs(H).max()
    assert np.allclose(E.element_energies(x).sum(), value, rtol=1e-14)


def brute_force_pattern(nv, tets):
    pairs = set()
    for cell in tets:
        dofs = [3 * v + c for v in cell for c in range(3)]
        for r in dofs:
            for c in dofs:
                pairs.add((c, r))
    pairs = sorted(pairs)  # column-major: sorted by column then row
    n = 3 * nv
    outer = np.zeros(n + 1, dtype=np.int32)
    for c, _ in pairs:
        outer[c + 1] += 1
    outer = np.cumsum(outer).astype(np.int32)
    inner = np.array([r for _, r in pairs], dtype=np.int32)
    return outer, inner


def test_pattern_and_slot_map_brute_force(oracle_mod):
    rng = np.random.default_rng(5)
    for n in (1, 2):
        X, tets = kuhn_grid(n)
        E = oracle_mod.OracleEnergy(len(X), tets, X, np.tile([[1.0, 1.0]], (len(tets), 1)))
        outer, inner = brute_force_pattern(len(X), tets)
        o, i = E.pattern()
        assert np.array_equal(o, outer) and np.array_equal(i, inner)
    # a mesh with an unused vertex and shuffled element order
    X, tets = kuhn_grid(2)
    perm = rng.permutation(len(tets))
    tets = tets[perm]
    Xp = np.vstack([X, [[5.0, 5.0, 5.0]]])
    E = oracle_mod.OracleEnergy(len(Xp), tets, Xp, np.tile([[1.0, 1.0]], (len(tets), 1)))
    outer, inner = brute_force_pattern(len(Xp), tets)
    o, i = E.pattern()
    assert np.array_equal(o, outer) and np.array_equal(i, inner)
    off = E.offsets()
    for e in range(len(tets)):
        dofs = [3 * v + c for v in tets[e] for c in range(3)]
        for j in range(12):
            col = dofs[j]
            for ii in range(12):
                pos = off[e, ii + 12 * j]
                assert outer[col] <= pos < outer[col + 1] and inner[pos] == dofs[ii]


@pytest.mark.parametrize("n", [2, 4, 10])
def test_kuhn_nnz_formula(oracle_mod, n):
    # nnz = 9 (V + 2 E_edges), E_edges = 3 n (n+1)^2 + 3 n^2 (n+1) + n^3 (axis, face-diagonal, body-diagonal edges)
    X, tets = kuhn_grid(n)
    E = oracle_mod.OracleEnergy(len(X), tets, X, np.tile([[1.0, 1.0]], (len(tets), 1)))
    edges = 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3
    assert E.nnz == 9 * ((n + 1) ** 3 + 2 * edges)


def test_assembly_is_bitwise_thread_invariant(oracle_mod):
    # mesh_energy_gtest.cpp:316-359, test_fem_batch_size.py:7-22
    X, tets = kuhn_grid(4)
    lam, mu = lame_parameters(1e5, 0.4)
    theta = np.tile([[float(lam), float(mu)]], (len(tets), 1))
    x = np.random.default_rng(2).uniform(-0.01, 0.01, X.size)
    E = oracle_mod.OracleEnergy(len(X), tets, X, theta)
    ref = E.eval(x, True, "projected", nthreads=1)
    for t in (2, 4, 8):
        got = E.eval(x, True, "projected", nthreads=t)
        assert got[0] == ref[0]
        assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])


def test_theta_linearity_and_parameter_derivatives(oracle_mod):
    X, tets = kuhn_grid(2)
    rng = np.random.default_rng(4)
    theta = np.stack([rng.uniform(1e4, 1e5, len(tets)), rng.uniform(1e4, 1e5, len(tets))], 1)
    x = rng.uniform(-0.02, 0.02, X.size)
    E = oracle_mod.OracleEnergy(len(X), tets, X, theta)
    value, g, _ = E.eval(x, True)
    dth = E.dE_dtheta(x)
    # psi is linear in theta: E = sum_e theta_e . dE/dtheta_e
    assert abs(np.sum(theta * dth) - value) <= 1e-12 * abs(value)
    # central differences in theta (theta override); E and g are linear in theta, so a large step has no
    # truncation error and little cancellation
    h = 1e3
    for e in (0, 7, 31):
        for p in range(2):
            tp, tm = theta.copy(), theta.copy()
            tp[e, p] += h
            tm[e, p] -= h
            fd = (E.eval(x, theta=tp)[0] - E.eval(x, theta=tm)[0]) / (2 * h)
            assert abs(fd - dth[e, p]) <= 1e-7 * (abs(dth[e, p]) + 1e-12)
    # force (= gradient) VJP / JVP in theta against differences of gradients
    w = rng.normal(size=x.size)
    vjp = E.force_theta_vjp(x, w)
    dtheta = rng.normal(size=theta.shape)
    jvp = E.force_theta_jvp(x, dtheta)
    assert abs(np.sum(vjp * dtheta) - w @ jvp) <= 1e-10 * (abs(w @ jvp) + 1e-30)
    gp = E.eval(x, True, theta=theta + h * dtheta)[1]
    gm = E.eval(x, True, theta=theta - h * dtheta)[1]
    assert np.abs((gp - gm) / (2 * h) - jvp).max() <= 1e-6 * np.abs(jvp).max()
    # the force at theta equals J theta exactly (force is linear in theta)
    assert np.abs(E.force_theta_jvp(x, theta) - g).max() <= 1e-12 * np.abs(g).max()
