"""Plastic parts (Fp != I) on the MI355X against the oracle, which restates MultiplicativeKinematics
(multiplicative_kinematics.cpp:9-51: Fe = F Fp^-1, W = det(Fp) psi(Fe), P and A pulled back) in its generic dense
route: energy, gradient, Exact and Projected Hessians, the differentiation products and the feasible step, each
within 1e-12 relative max-norm; and the evaluation-time errors of a rejected Fp."""
import numpy as np
import pytest
import scipy.sparse as sp

from litepgo_amd._native import DomainError
from litepgo_amd import VolumetricFemEnergy, kuhn_grid, lame_parameters

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module", autouse=True)
def torch():
    # Initialise torch's HIP context before the library's first call in this process.
    import torch

    assert torch.cuda.is_available()
    return torch


@pytest.fixture(scope="module")
def case(oracle_mod):
    X, tets = kuhn_grid(6)
    rng = np.random.default_rng(4)
    T = tets.shape[0]
    E = np.exp(rng.normal(np.log(1e5), 0.3, size=T))
    lam, mu = lame_parameters(E, 0.4)
    theta = np.stack([lam, mu], axis=1)
    Fp = np.eye(3)[None] + 0.15 * rng.normal(size=(T, 3, 3))  # generic plastic parts, det > 0
    assert np.all(np.linalg.det(Fp) > 0)
    u = 0.03 * rng.normal(size=X.size) / 6
    dev = VolumetricFemEnergy(len(X), tets, X, theta, Fp=Fp)
    ref = oracle_mod.OracleEnergy(len(X), tets, X, theta, Fp=Fp)
    return X, tets, theta, Fp, u, dev, ref


@pytest.mark.parametrize("kind", ["exact", "projected"])
def test_energy_gradient_hessian(case, kind):
    X, tets, theta, Fp, u, dev, ref = case
    v, g, H = dev.eval_values(u, True, kind)
    vr, gr, Hr = ref.eval(u, True, kind)
    print(f"{kind}: dE {abs(v - vr) / abs(vr):.2e} dg {rel(g, gr):.2e} dH {rel(H, Hr):.2e}", flush=True)
    assert abs(v - vr) <= TOL * abs(vr)
    assert rel(g, gr) <= TOL
    assert rel(H, Hr) <= TOL
    if kind == "projected":
        assert dev.last_indefinite_count > 0  # the plastic parts make some tangents indefinite


def test_derivative_products(torch, case):
    X, tets, theta, Fp, u, dev, ref = case
    x = torch.from_numpy(u).cuda()
    rng = np.random.default_rng(5)
    assert rel(dev.dE_dtheta_device(x).cpu().numpy(), ref.dE_dtheta(u)) <= TOL
    w = rng.normal(size=u.size)
    assert rel(dev.force_vjp_theta_device(x, torch.from_numpy(w).cuda()).cpu().numpy(),
               ref.force_theta_vjp(u, w)) <= TOL
    dth = rng.normal(size=theta.shape)
    assert rel(dev.force_jvp_theta_device(x, torch.from_numpy(dth).cuda()).cpu().numpy(),
               ref.force_theta_jvp(u, dth)) <= TOL
    _, _, Hr = ref.eval(u, False, "exact")
    outer, inner = ref.pattern()
    Hm = sp.csc_matrix((Hr, inner, outer), shape=(ref.num_dofs, ref.num_dofs))
    vv = rng.normal(size=u.size)
    assert rel(dev.hvp_device(x, torch.from_numpy(vv).cuda()).cpu().numpy(), Hm @ vv) <= TOL


def test_feasible_step(case):
    X, tets, theta, Fp, u, dev, ref = case
    rng = np.random.default_rng(6)
    for scale in (0.01, 0.2, 2.0):
        dx = rng.normal(scale=scale, size=u.size)
        a, r = dev.max_feasible_step(u, dx), ref.max_feasible_step(u, dx)
        assert (a == r == 1.0) or abs(a - r) <= TOL * r


def test_rejected_plastic_parts_fail_evaluations(torch):
    X, tets = kuhn_grid(2)
    T = tets.shape[0]
    lam, mu = lame_parameters(1e5, 0.4)
    theta = np.tile([lam, mu], (T, 1))
    u = np.zeros(X.size)
    for Fp_bad, kind in ((-1.0, "det"), (np.nan, "finite")):
        Fp = np.stack([np.eye(3)] * T)
        Fp[3] = Fp_bad * np.eye(3) if kind == "det" else np.full((3, 3), np.nan)
        dev = VolumetricFemEnergy(len(X), tets, X, theta, Fp=Fp)
        with pytest.raises(ValueError):
            dev.eval(u, True, "exact")
        with pytest.raises(ValueError):
            dev.max_feasible_step(u, u)


def colmajor(Fp):
    return np.ascontiguousarray(np.asarray(Fp).transpose(0, 2, 1).reshape(len(Fp), 9))


def test_force_fp_response(torch, case, oracle_mod):
    """VolumetricFemForceFpResponse (volumetric_fem_force_response.hpp:81-179, neo_hookean.cpp:103-128): value,
    forward and transpose products in the per-tet column-major vec(Fp) against the oracle, with the handle's own Fp
    and with per-call plastic parts that differ from it."""
    X, tets, theta, Fp, u, dev, ref = case
    T = tets.shape[0]
    x = torch.from_numpy(u).cuda()
    _, g, _ = ref.eval(u, True)
    assert rel(dev.force_fp_device(x).cpu().numpy(), g) <= TOL
    assert rel(dev.force_fp_device(x, torch.from_numpy(colmajor(Fp)).cuda()).cpu().numpy(), g) <= TOL
    rng = np.random.default_rng(8)
    Fp2 = np.eye(3)[None] + 0.1 * rng.normal(size=(T, 3, 3))
    assert np.all(np.linalg.det(Fp2) > 0)
    fp2 = colmajor(Fp2)
    w = rng.normal(size=u.size)
    d = rng.normal(size=(T, 9))
    fr, jtr, jvr = oracle_mod.force_fp_products(tets, X, theta, u, fp2, w=w, dFp=d)
    fpd = torch.from_numpy(fp2).cuda()
    f = dev.force_fp_device(x, fpd).cpu().numpy()
    jt = dev.force_vjp_fp_device(x, torch.from_numpy(w).cuda(), fpd).cpu().numpy()
    jv = dev.force_jvp_fp_device(x, torch.from_numpy(d).cuda(), fpd).cpu().numpy()
    print(f"force_fp: df {rel(f, fr):.2e} dJtw {rel(jt, jtr):.2e} dJd {rel(jv, jvr):.2e}", flush=True)
    assert rel(f, fr) <= TOL and rel(jt, jtr) <= TOL and rel(jv, jvr) <= TOL
    # <w, J d> = <J^T w, d> between the two device kernels.
    assert abs(w @ jv - np.sum(jt * d)) <= 1e-12 * (np.abs(w) @ np.abs(jv))
    # A per-call theta overrides the snapshot in all three products.
    theta2 = theta * np.array([1.3, 0.8])
    td = torch.from_numpy(theta2).cuda()
    fr2, jtr2, jvr2 = oracle_mod.force_fp_products(tets, X, theta2, u, fp2, w=w, dFp=d)
    assert rel(dev.force_fp_device(x, fpd, theta=td).cpu().numpy(), fr2) <= TOL
    assert rel(dev.force_vjp_fp_device(x, torch.from_numpy(w).cuda(), fpd, theta=td).cpu().numpy(), jtr2) <= TOL
    assert rel(dev.force_jvp_fp_device(x, torch.from_numpy(d).cuda(), fpd, theta=td).cpu().numpy(), jvr2) <= TOL
    # A plain handle (every Fp = I) takes per-call plastic parts the same way.
    plain = VolumetricFemEnergy(len(X), tets, X, theta)
    assert rel(plain.force_fp_device(x, fpd).cpu().numpy(), fr) <= TOL
    assert rel(plain.force_fp_device(x).cpu().numpy(), plain.eval(u, True)[1]) <= TOL


def test_force_fp_response_errors_and_unused_vertex(torch, oracle_mod):
    X, tets = kuhn_grid(2)
    X = np.vstack([X, [[5.0, 5.0, 5.0]]])  # a vertex no tet uses: zero force and zero products
    T = tets.shape[0]
    lam, mu = lame_parameters(1e5, 0.4)
    theta = np.tile([lam, mu], (T, 1))
    rng = np.random.default_rng(9)
    u = 0.01 * rng.normal(size=X.size)
    dev = VolumetricFemEnergy(len(X), tets, X, theta)
    Fp = np.eye(3)[None] + 0.1 * rng.normal(size=(T, 3, 3))
    fp = colmajor(Fp)
    x = torch.from_numpy(u).cuda()
    d = rng.normal(size=(T, 9))
    f = dev.force_fp_device(x, torch.from_numpy(fp).cuda()).cpu().numpy()
    jv = dev.force_jvp_fp_device(x, torch.from_numpy(d).cuda(), torch.from_numpy(fp).cuda()).cpu().numpy()
    fr, _, jvr = oracle_mod.force_fp_products(tets, X, theta, u, fp, dFp=d)
    assert rel(f, fr) <= TOL and rel(jv, jvr) <= TOL
    assert np.all(f[-3:] == 0.0) and np.all(jv[-3:] == 0.0)
    # MultiplicativeKinematics' two rejections (multiplicative_kinematics.cpp:9-16): a non-finite Fp is an
    # invalid_argument, a non-positive det Fp a domain_error, each with the reference's message.
    bad = fp.copy()
    bad[5] = (-np.eye(3)).ravel()
    for call in (lambda b: dev.force_fp_device(x, b), lambda b: dev.force_vjp_fp_device(x, x, b),
                 lambda b: dev.force_jvp_fp_device(x, torch.from_numpy(d).cuda(), b)):
        with pytest.raises(DomainError, match="MultiplicativeKinematics requires finite det"):
            call(torch.from_numpy(bad).cuda())
    nanfp = fp.copy()
    nanfp[7, 4] = np.nan
    with pytest.raises(ValueError, match="MultiplicativeKinematics requires finite F and Fp") as info:
        dev.force_fp_device(x, torch.from_numpy(nanfp).cuda())
    assert not isinstance(info.value, DomainError)
    # The handle recovers after the error.
    assert rel(dev.force_fp_device(x, torch.from_numpy(fp).cuda()).cpu().numpy(), fr) <= TOL
    # A handle whose snapshot Fp was rejected still answers calls that bring their own plastic parts.
    snap = np.stack([np.eye(3)] * T)
    snap[2] = -np.eye(3)
    rejected = VolumetricFemEnergy(len(X), tets, X, theta, Fp=snap)
    with pytest.raises(DomainError):
        rejected.force_fp_device(x)
    assert rel(rejected.force_fp_device(x, torch.from_numpy(fp).cuda()).cpu().numpy(), fr) <= TOL


def test_force_fp_fits_a_plastic_scale(torch):
    """The reference's Gauss-Newton fit of a shared plastic scale (tests/python/test_fp_force_response.py:44-66): the
    force at Fp = 1.05 I is the target, and Gauss-Newton on the scalar s (residual f(s I) - target, Jacobian the
    forward product along I in every tet) recovers s = 1.05 from 1."""
    X, tets = kuhn_grid(3)
    T = tets.shape[0]
    lam, mu = lame_parameters(1e5, 0.4)
    dev = VolumetricFemEnergy(len(X), tets, X, np.tile([lam, mu], (T, 1)))
    rng = np.random.default_rng(12)
    x = torch.from_numpy(0.02 * rng.normal(size=X.size)).cuda()
    direction = torch.from_numpy(np.tile(np.eye(3).ravel(order="F"), (T, 1))).cuda()
    target = dev.force_fp_device(x, 1.05 * direction)
    s = 1.0
    for _ in range(20):
        r = dev.force_fp_device(x, s * direction) - target
        Jd = dev.force_jvp_fp_device(x, direction, s * direction)
        grad = float(torch.dot(Jd, r))
        if abs(grad) <= 1e-10 * float(torch.linalg.vector_norm(Jd)) * float(torch.linalg.vector_norm(target)):
            break
        s -= grad / float(torch.dot(Jd, Jd))
    assert abs(s - 1.05) <= 1e-8


def test_force_fp_closed_form_dilation(torch):
    """One unit tet at F = a I with Fp = b I (the reference's analytic case, neo_hookean_gtest.cpp:270-281): the force
    is f_b = w b^2 (mu a/b + (3 lambda log(a/b) - mu) b/a) g_b, the forward product along vec(I) is
    w (mu a + b^2/a (9 lambda log(a/b) - 3 mu - 3 lambda)) g_b, and the transpose product of any w_f dotted with
    vec(I) equals w_f . (J vec(I)); w = 1/6, g_0 = -(1, 1, 1), g_i = e_i."""
    lam, mu, a, b = 2.0, 3.0, 1.3, 0.8
    X = np.vstack([np.zeros(3), np.eye(3)])
    dev = VolumetricFemEnergy(4, np.array([[0, 1, 2, 3]], dtype=np.int32), X, np.array([[lam, mu]]))
    G = np.vstack([-np.ones(3), np.eye(3)])
    x = torch.from_numpy(((a - 1.0) * X).ravel()).cuda()
    vecI = np.eye(3).ravel(order="F")[None]
    fp = torch.from_numpy(b * vecI).cuda()
    f = dev.force_fp_device(x, fp).cpu().numpy()
    f_expected = (b * b * (mu * a / b + (3 * lam * np.log(a / b) - mu) * b / a) / 6.0) * G.ravel()
    np.testing.assert_allclose(f, f_expected, rtol=1e-13, atol=1e-14)
    jv = dev.force_jvp_fp_device(x, torch.from_numpy(vecI).cuda(), fp).cpu().numpy()
    scalar = mu * a + b * b / a * (9 * lam * np.log(a / b) - 3 * mu - 3 * lam)
    np.testing.assert_allclose(jv, scalar / 6.0 * G.ravel(), rtol=1e-13, atol=1e-14)
    wf = np.random.default_rng(13).normal(size=12)
    jt = dev.force_vjp_fp_device(x, torch.from_numpy(wf).cuda(), fp).cpu().numpy()
    assert abs(float(jt.ravel() @ vecI.ravel()) - float(wf @ jv)) <= 1e-13 * np.abs(wf) @ np.abs(jv)


def test_force_fp_on_the_reference_box_mesh(torch, oracle_mod):
    """The reference's box-tet.veg asset (827 tets, irregular numbering) with generic plastic parts and per-tet
    material: value, forward and transpose products against the oracle."""
    from pathlib import Path

    from litepgo_amd import read_veg

    mesh = read_veg(Path(__file__).resolve().parent / "golden" / "box-tet.veg")
    T = len(mesh.cells)
    rng = np.random.default_rng(14)
    E = np.exp(rng.normal(np.log(1e5), 0.3, size=T))
    lam, mu = lame_parameters(E, 0.4)
    theta = np.stack([lam, mu], axis=1)
    dev = VolumetricFemEnergy(mesh.num_vertices, mesh.cells, mesh.positions, theta)
    u = 1e-3 * rng.normal(size=3 * mesh.num_vertices)
    Fp = np.eye(3)[None] + 0.1 * rng.normal(size=(T, 3, 3))
    fp = colmajor(Fp)
    w = rng.normal(size=u.size)
    d = rng.normal(size=(T, 9))
    fr, jtr, jvr = oracle_mod.force_fp_products(mesh.cells, mesh.positions, theta, u, fp, w=w, dFp=d)
    x = torch.from_numpy(u).cuda()
    fpd = torch.from_numpy(fp).cuda()
    assert rel(dev.force_fp_device(x, fpd).cpu().numpy(), fr) <= TOL
    assert rel(dev.force_vjp_fp_device(x, torch.from_numpy(w).cuda(), fpd).cpu().numpy(), jtr) <= TOL
    assert rel(dev.force_jvp_fp_device(x, torch.from_numpy(d).cuda(), fpd).cpu().numpy(), jvr) <= TOL
