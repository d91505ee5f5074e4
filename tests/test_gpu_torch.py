"""The torch bridge (litepgo_amd.torch) on the MI355X: reverse- and forward-mode derivatives of the energy and
the force with respect to positions and per-tet material, checked by central finite differences (step 1e-6) on
random directions, including the 32^3 differentiation workload (per-tet log-normal E, nu ~ U[0.3, 0.45]) on 64
directions; CPU tensors round-trip; double backward gives Hessian-vector products."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEP = 1e-6


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available()
    return torch


def setup(torch, w, E, nu):
    from litepgo_amd import VolumetricFemEnergy

    pot = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
    x = torch.from_numpy(w.u).cuda().requires_grad_(True)
    Et = torch.from_numpy(np.asarray(E, dtype=np.float64)).cuda().requires_grad_(True)
    nut = torch.from_numpy(np.asarray(nu, dtype=np.float64)).cuda().requires_grad_(True)
    return pot, x, Et, nut


def central(fn, h):
    return (fn(h) - fn(-h)) / (2 * h)


def relerr(a, b, scale):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / scale)


def check_all(torch, pot, x, Et, nut, ndir, seed):
    from litepgo_amd import torch as lt

    rng = np.random.default_rng(seed)
    theta = lt.lame(Et, nut)
    E, f = lt.energy_and_force(pot, x, theta)
    w = torch.from_numpy(rng.normal(size=pot.num_dofs)).cuda()
    # Reverse mode: gradients of E + w . f with respect to x, E_young and nu.
    gx, gE, gnu = torch.autograd.grad(E + (w * f).sum(), (x, Et, nut))
    x0, E0, nu0 = x.detach(), Et.detach(), nut.detach()

    def scalar(dx, dE, dnu):
        th = lt.lame(E0 + dE, nu0 + dnu)
        e, ff = lt.energy_and_force(pot, x0 + dx, th)
        return float(e) + float((w * ff).sum())

    # Each directional derivative is checked relative to its own size: |fd - ad| <= 1e-6 |ad| + floor, with the floor
    # the rounding error of a central difference of the scalar objective L (4 eps |L| / step).
    floor = 4 * np.finfo(np.float64).eps * abs(scalar(0.0, 0.0, 0.0)) / STEP
    worst = {"x": 0.0, "E,nu": 0.0}  # largest |fd - ad| / |ad|
    ratio = 0.0                      # largest |fd - ad| / (1e-6 |ad| + floor); must stay below 1
    for k in range(ndir):
        d = torch.from_numpy(rng.normal(size=pot.num_dofs)).cuda()
        d = d / torch.linalg.vector_norm(d)
        fd = central(lambda h: scalar(h * d, 0.0, 0.0), STEP)
        ad = float((gx * d).sum())
        worst["x"] = max(worst["x"], abs(fd - ad) / abs(ad))
        ratio = max(ratio, abs(fd - ad) / (1e-6 * abs(ad) + floor))
        dE = torch.from_numpy(rng.normal(size=pot.num_tets)).cuda() * E0
        dnu = torch.from_numpy(rng.normal(size=pot.num_tets)).cuda() * 0.01
        fd = central(lambda h: scalar(0.0, h * dE, h * dnu), STEP)
        ad = float((gE * dE).sum() + (gnu * dnu).sum())
        worst["E,nu"] = max(worst["E,nu"], abs(fd - ad) / abs(ad))
        ratio = max(ratio, abs(fd - ad) / (1e-6 * abs(ad) + floor))
    print("worst relative FD mismatch", worst, "floor", floor, "ratio to tolerance", ratio, flush=True)
    assert ratio < 1.0


def test_small_mesh_derivatives(torch):
    from litepgo_amd import workloads

    w, E, nu = workloads.differentiate(seed=1, n=4)
    pot, x, Et, nut = setup(torch, w, E, nu)
    check_all(torch, pot, x, Et, nut, ndir=16, seed=2)


def test_differentiation_workload_32cube_64_directions(torch):
    from litepgo_amd import workloads

    w, E, nu = workloads.differentiate(seed=0, n=32)
    pot, x, Et, nut = setup(torch, w, E, nu)
    check_all(torch, pot, x, Et, nut, ndir=64, seed=3)


def test_hessian_vector_product_by_double_backward_and_forward_mode(torch):
    import torch.autograd.forward_ad as fwAD

    from litepgo_amd import torch as lt
    from litepgo_amd import workloads

    w, E, nu = workloads.differentiate(seed=4, n=6)
    pot, x, Et, nut = setup(torch, w, E, nu)
    theta = lt.lame(Et, nut).detach()
    rng = np.random.default_rng(5)
    v = torch.from_numpy(rng.normal(size=pot.num_dofs)).cuda()
    E_ = lt.energy(pot, x, theta)
    (g,) = torch.autograd.grad(E_, x, create_graph=True)
    (hv,) = torch.autograd.grad((g * v).sum(), x)
    x0 = x.detach()
    fd = (lt.force(pot, x0 + STEP * v, theta) - lt.force(pot, x0 - STEP * v, theta)) / (2 * STEP)
    scale = float(fd.abs().max())
    assert relerr(hv.cpu(), fd.cpu(), scale) < 1e-6
    with fwAD.dual_level():
        dual = fwAD.make_dual(x0, v)
        tangent = fwAD.unpack_dual(lt.force(pot, dual, theta)).tangent
    assert relerr(tangent.cpu(), hv.cpu(), scale) < 1e-12
    # dE/dtheta through double backward in x: d/dx (u . dE/dtheta) = (df/dtheta) u
    th = theta.clone().requires_grad_(True)
    E2 = lt.energy(pot, x, th)
    (gth,) = torch.autograd.grad(E2, th, create_graph=True)
    # f is linear in theta, so central FD is exact up to rounding; scale u with theta so the step is relative.
    u = torch.from_numpy(rng.normal(size=(pot.num_tets, 2))).cuda() * theta
    (mix,) = torch.autograd.grad((gth * u).sum(), x)
    fd = (lt.force(pot, x0, theta + STEP * u) - lt.force(pot, x0, theta - STEP * u)) / (2 * STEP)
    assert relerr(mix.cpu(), fd.cpu(), float(fd.abs().max())) < 1e-6


def test_cpu_tensors_round_trip(torch):
    from litepgo_amd import VolumetricFemEnergy, workloads
    from litepgo_amd import torch as lt

    w = workloads.plain_small(seed=2)
    pot = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
    x = torch.from_numpy(w.u).requires_grad_(True)  # CPU, as the reference's bridge requires
    E = lt.energy(pot, x)
    E.backward()
    assert E.device.type == "cpu" and x.grad.device.type == "cpu"
    v, g, _ = pot.eval_values(w.u, True, None)
    assert float(E) == v and np.array_equal(x.grad.numpy(), g)


def test_lame_kernels(torch):
    from litepgo_amd import torch as lt
    from litepgo_amd.fem import lame_parameters

    rng = np.random.default_rng(6)
    E = np.exp(rng.normal(np.log(1e5), 0.5, size=1001))
    nu = rng.uniform(0.3, 0.45, size=1001)
    Et = torch.from_numpy(E).cuda().requires_grad_(True)
    nut = torch.from_numpy(nu).cuda().requires_grad_(True)
    th = lt.lame(Et, nut)
    lam, mu = lame_parameters(E, nu)
    # Same operations in the same order, IEEE division: bit-identical.
    assert np.array_equal(th.detach().cpu().numpy(), np.stack([lam, mu], axis=1))
    g = torch.from_numpy(rng.normal(size=(1001, 2))).cuda()
    gE, gnu = torch.autograd.grad((th * g).sum(), (Et, nut))
    # Reverse product against autograd of the same formula in torch (float64, CPU).
    Ec = torch.from_numpy(E).requires_grad_(True)
    nuc = torch.from_numpy(nu).requires_grad_(True)
    ref = torch.stack([Ec * nuc / ((1.0 + nuc) * (1.0 - 2.0 * nuc)), Ec / (2.0 * (1.0 + nuc))], dim=-1)
    rE, rnu = torch.autograd.grad((ref * g.cpu()).sum(), (Ec, nuc))
    assert relerr(gE.cpu(), rE, float(rE.abs().max())) < 1e-14
    assert relerr(gnu.cpu(), rnu, float(rnu.abs().max())) < 1e-14


def test_fused_vjp_matches_separate_products(torch):
    from litepgo_amd import VolumetricFemEnergy, workloads

    w, E, nu = workloads.differentiate(seed=7, n=8)
    pot = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
    rng = np.random.default_rng(8)
    x = torch.from_numpy(w.u).cuda()
    theta = torch.from_numpy(w.theta * rng.uniform(0.9, 1.1, size=w.theta.shape)).cuda().contiguous()
    wv = torch.from_numpy(rng.normal(size=pot.num_dofs)).cuda()
    s = torch.tensor([0.7], dtype=torch.float64, device="cuda")
    gx, gt = pot.vjp_device(x, gE=s, w=wv, theta=theta)
    _, f, _ = pot.eval_device(x, True, None, theta=theta)
    ref_x = 0.7 * f + pot.hvp_device(x, wv, theta)
    ref_t = 0.7 * pot.dE_dtheta_device(x, theta) + pot.force_vjp_theta_device(x, wv, theta)
    assert relerr(gx.cpu(), ref_x.cpu(), float(ref_x.abs().max())) < 1e-12
    assert relerr(gt.cpu(), ref_t.cpu(), float(ref_t.abs().max())) < 1e-12
    # Either cotangent alone, and either output alone.
    gx1, _ = pot.vjp_device(x, gE=s, theta=theta, want_theta=False)
    assert relerr(gx1.cpu(), (0.7 * f).cpu(), float(f.abs().max())) < 1e-12
    _, gt1 = pot.vjp_device(x, w=wv, theta=theta, want_x=False)
    assert torch.equal(gt1, pot.force_vjp_theta_device(x, wv, theta))
    gx0, gt0 = pot.vjp_device(x, theta=theta)
    assert float(gx0.abs().max()) == 0.0 and float(gt0.abs().max()) == 0.0
    # Energy written to a device tensor equals the returned value.
    Ed, _, _ = pot.eval_device_energy(x, False, None, theta=theta)
    v, _, _ = pot.eval_device(x, False, None, theta=theta)
    assert float(Ed[0]) == v


def test_force_fp_reverse_and_forward_modes(torch):
    """force_fp: autograd's reverse product against central differences of the force in Fp, and forward-mode AD
    against the same differences (the reference's Fp force response, tests/python/test_fp_force_response.py,
    differentiates the same map)."""
    from litepgo_amd import VolumetricFemEnergy, kuhn_grid, lame_parameters
    from litepgo_amd import torch as lt

    X, tets = kuhn_grid(3)
    T = tets.shape[0]
    rng = np.random.default_rng(11)
    lam, mu = lame_parameters(1e5, 0.4)
    pot = VolumetricFemEnergy(len(X), tets, X, np.tile([lam, mu], (T, 1)))
    x = torch.from_numpy(0.02 * rng.normal(size=X.size)).cuda()
    Fp0 = np.eye(3)[None] + 0.1 * rng.normal(size=(T, 3, 3))
    Fp = torch.from_numpy(np.ascontiguousarray(Fp0.transpose(0, 2, 1).reshape(T, 9))).cuda().requires_grad_(True)
    f = lt.force_fp(pot, Fp, x)
    w = torch.from_numpy(rng.normal(size=pot.num_dofs)).cuda()
    (gFp,) = torch.autograd.grad((w * f).sum(), (Fp,))
    d = torch.from_numpy(rng.normal(size=(T, 9))).cuda()
    F0 = Fp.detach()
    fd = central(lambda h: lt.force_fp(pot, F0 + h * d, x), STEP)
    ad = float((gFp * d).sum())
    assert abs(float((w * fd).sum()) - ad) <= 1e-6 * (abs(ad) + float(torch.linalg.vector_norm(gFp)))
    import torch.autograd.forward_ad as fwAD

    with fwAD.dual_level():
        tangent = fwAD.unpack_dual(lt.force_fp(pot, fwAD.make_dual(F0, d), x)).tangent
    scale = float(torch.abs(fd).max())
    assert float(torch.abs(tangent - fd).max()) <= 1e-6 * scale


def test_force_fp_rejects_gradients_in_x_and_theta(torch):
    """force_fp differentiates in Fp only: asking autograd for df/dx or df/dtheta through it raises instead of
    returning zeros."""
    from litepgo_amd import VolumetricFemEnergy, kuhn_grid, lame_parameters
    from litepgo_amd import torch as lt

    X, tets = kuhn_grid(2)
    T = tets.shape[0]
    lam, mu = lame_parameters(1e5, 0.4)
    pot = VolumetricFemEnergy(len(X), tets, X, np.tile([lam, mu], (T, 1)))
    rng = np.random.default_rng(4)
    Fp = torch.from_numpy(np.tile(np.eye(3).reshape(1, 9), (T, 1))).cuda().requires_grad_(True)
    x = torch.from_numpy(0.01 * rng.normal(size=X.size)).cuda().requires_grad_(True)
    f = lt.force_fp(pot, Fp, x)
    with pytest.raises(RuntimeError, match="only Fp is differentiable"):
        torch.autograd.grad(f.sum(), (x,))
    theta = torch.from_numpy(np.tile([lam, mu], (T, 1))).cuda().requires_grad_(True)
    f = lt.force_fp(pot, Fp, x.detach(), theta)
    with pytest.raises(RuntimeError, match="only Fp is differentiable"):
        torch.autograd.grad(f.sum(), (theta,))
    # Fp alone still works.
    f = lt.force_fp(pot, Fp, x.detach())
    (g,) = torch.autograd.grad(f.sum(), (Fp,))
    assert g.shape == Fp.shape and torch.isfinite(g).all()
