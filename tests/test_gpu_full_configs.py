"""The two largest benchmark configurations checked at their full size on the MI355X.

Batched: 256 Kuhn meshes with per-axis counts U{4..24} (h = 0.1, per-mesh E log-uniform in [1e4, 1e6],
nu U[0.3, 0.48], noise +-0.05 h; benchmarks/mesh_batch_benchmark.py:11-35 times the reference's per-mesh path on
such a set).  Every mesh of the batch equals its own plain evaluation bit for bit, before and after the launch
autotuner runs; eight sampled meshes, the largest among them, agree with the oracle to 1e-12; every mesh's forces
sum to zero and its energy, gradient and Hessian are unchanged (to rounding) by a rigid translation.

Equilibrium: the 48^3 cantilever (x = 0 face fixed, gravity 9.8 along -z, density 1000, per-tet E log-normal) and
the 8-beam set with 40..48 cells per axis (tests/cpp/solver/static_equilibrium_response_gtest.cpp:14-89 checks
the reference's equilibrium response the same ways on a small problem): convergence to |g|_inf < 1e-8, the free
DOFs' internal forces equal to the gravity loads and the support reactions carrying the weight, the transpose
identity w . (J v) = (J^T w) . v at full size, two central finite differences of re-solved equilibria against the
adjoint, and the 8 beams solved as one union against separate solves.
"""
import numpy as np
import pytest

from litepgo_amd import BatchedFemEnergy, VolumetricFemEnergy, solver, workloads

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    s = np.abs(b).max() if b.size else 0.0
    return float(np.abs(a - b).max() / s) if s else float(np.abs(a).max() if a.size else 0.0)


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available()
    return torch


# ---------------------------------------------------------------------------------------------------------------
# Batched configuration: 256 meshes, 4..24 cells per axis.
# ---------------------------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def batch256(torch):
    meshes = workloads.batched(256)
    batch = BatchedFemEnergy(meshes)
    x = torch.from_numpy(np.concatenate([w.u for w in meshes])).cuda()
    print(f"batch of {len(meshes)} meshes, {batch.num_tets} tets, {batch.nnz} Hessian values", flush=True)
    return meshes, batch, x


def _plain_matches(torch, meshes, batch, E, g, H, kind, label):
    vo, _, no = batch.offsets()
    for m, w in enumerate(meshes):
        plain = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
        xm = torch.from_numpy(w.u).cuda()
        v, gp, Hp = plain.eval_device(xm, True, kind)
        assert E[m].item() == v, (label, m, w.name, E[m].item(), v)
        assert torch.equal(g[3 * int(vo[m]):3 * int(vo[m + 1])], gp), (label, m, w.name)
        assert torch.equal(H[int(no[m]):int(no[m + 1])], Hp), (label, m, w.name)


@pytest.mark.parametrize("kind", ["exact", "projected"])
def test_batch256_equals_plain_bitwise(torch, batch256, kind):
    meshes, batch, x = batch256
    E, g, H = batch.eval_device(x, True, kind)
    print(f"{kind}: comparing 256 meshes with their plain evaluations", flush=True)
    _plain_matches(torch, meshes, batch, E, g, H, kind, "default launch")


def test_batch256_autotuned_equals_plain_bitwise(torch, batch256):
    """The autotuner's choice changes no bit of any mesh (assembly sums in an order fixed by the mesh alone)."""
    meshes, batch, x = batch256
    shape, ms = batch.autotune(x, "projected", repeats=2)
    print(f"autotuned shape {shape} {ms:.3f} ms", flush=True)
    E, g, H = batch.eval_device(x, True, "projected")
    _plain_matches(torch, meshes, batch, E, g, H, "projected", f"autotuned {shape}")


@pytest.mark.parametrize("kind", ["exact", "projected"])
def test_batch256_sampled_meshes_match_oracle(torch, batch256, oracle_mod, kind):
    meshes, batch, x = batch256
    E, g, H = batch.eval_device(x, True, kind)
    E, g, H = E.cpu().numpy(), g.cpu().numpy(), H.cpu().numpy()
    vo, _, no = batch.offsets()
    sizes = np.array([w.tets.shape[0] for w in meshes])
    picks = sorted(set([int(np.argmax(sizes)), int(np.argmin(sizes))] +
                       list(np.random.default_rng(7).choice(len(meshes), 6, replace=False))))
    for m in picks:
        w = meshes[m]
        print(f"{kind}: oracle on mesh {m} {w.name} ({w.tets.shape[0]} tets)", flush=True)
        ref = oracle_mod.OracleEnergy(w.num_vertices, w.tets, w.rest, w.theta)
        vr, gr, Hr = ref.eval(w.u, True, kind)
        assert abs(E[m] - vr) <= TOL * abs(vr), (m, E[m], vr)
        assert rel(g[3 * vo[m]:3 * vo[m + 1]], gr) <= TOL, m
        assert rel(H[no[m]:no[m + 1]], Hr) <= TOL, m


def test_batch256_translation_invariance_and_zero_net_force(torch, batch256):
    meshes, batch, x = batch256
    E0, g0, H0 = (t.cpu().numpy() for t in batch.eval_device(x, True, "exact"))
    vo, _, no = batch.offsets()
    rng = np.random.default_rng(3)
    shift = np.concatenate([np.tile(rng.uniform(-0.3, 0.3, size=3), w.num_vertices) for w in meshes])
    E1, g1, H1 = (t.cpu().numpy() for t in batch.eval_device(x + torch.from_numpy(shift).cuda(), True, "exact"))
    for m in range(len(meshes)):
        gm = g0[3 * vo[m]:3 * vo[m + 1]].reshape(-1, 3)
        # Sum of the nodal forces: zero up to the rounding of the per-vertex sums.
        net = np.abs(gm.sum(axis=0)).max()
        assert net <= 1e-11 * np.abs(gm).sum(), (m, net)
        assert abs(E1[m] - E0[m]) <= 1e-10 * abs(E0[m]), m
        assert rel(g1[3 * vo[m]:3 * vo[m + 1]], g0[3 * vo[m]:3 * vo[m + 1]]) <= 1e-9, m
        assert rel(H1[no[m]:no[m + 1]], H0[no[m]:no[m + 1]]) <= 1e-9, m


# ---------------------------------------------------------------------------------------------------------------
# Equilibrium configuration: 48^3 cantilever and 8 beams of 40..48 cells per axis.
# ---------------------------------------------------------------------------------------------------------------
def _opts(tol=1e-8):
    return solver.NewtonOptions(gradient_tolerance=tol)


@pytest.fixture(scope="module")
def beam48(torch):
    b = workloads.beam(48, seed=0)
    dev = VolumetricFemEnergy(b.num_vertices, b.tets, b.rest, b.theta)
    lin = torch.from_numpy(-b.load).cuda()
    x0 = torch.zeros(dev.num_dofs, dtype=torch.float64, device="cuda")
    res = solver.newton_solve(dev, b.fixed, x0, _opts(), linear=lin)
    print(f"48^3: status {res.status} newton {res.iterations} pcg {res.pcg_iterations} |g| {res.gradient_norm:.3g}",
          flush=True)
    return b, dev, lin, res


def test_beam48_converges_and_balances_gravity(torch, beam48):
    b, dev, lin, res = beam48
    assert res.status == solver.NewtonStatus.Converged and res.gradient_norm < 1e-8
    x = res.x.cpu().numpy()
    assert np.all(x[b.fixed] == 0.0)
    _, f, _ = dev.eval_device(res.x, True, None)
    f = f.cpu().numpy()
    free = ~b.fixed
    # Free DOFs: internal force = applied gravity load.
    assert np.abs(f[free] - b.load[free]).max() < 1e-8
    # The total load is density * g * volume = 9800 along -z.  The internal forces sum to zero over all DOFs
    # (translation invariance), so the supports' internal forces balance the loads on the free vertices: they
    # carry the weight.
    load = b.load.reshape(-1, 3)
    assert abs(load[:, 2].sum() + 9800.0) <= 1e-9 * 9800.0
    fv = f.reshape(-1, 3)
    on_support = b.fixed.reshape(-1, 3)[:, 0]
    assert np.abs(fv.sum(axis=0)).max() <= 1e-9 * 9800.0
    support = fv[on_support].sum(axis=0)
    carried = -load[~on_support].sum(axis=0)
    print(f"support forces {support} free-vertex weight {carried}", flush=True)
    assert np.abs(support - carried).max() <= 1e-6 * 9800.0


def test_beam48_transpose_identity_and_finite_differences(torch, beam48):
    b, dev, lin, res = beam48
    rng = np.random.default_rng(5)
    n = dev.num_dofs
    w = rng.normal(size=n)
    v = rng.normal(size=b.theta.shape) * b.theta
    opts = solver.NewtonOptions(gradient_tolerance=1e-8, pcg_rtol=1e-12)
    gw = solver.equilibrium_vjp(dev, b.fixed, res.x, torch.from_numpy(w).cuda(), opts).cpu().numpy()
    jv = solver.equilibrium_jvp(dev, b.fixed, res.x, torch.from_numpy(v).cuda(), opts).cpu().numpy()
    lhs, rhs = float(w @ jv), float((gw * v).sum())
    print(f"w.(Jv) {lhs!r} (J^T w).v {rhs!r}", flush=True)
    assert abs(lhs - rhs) <= 1e-8 * abs(lhs)
    # Central differences of re-solved equilibria along two theta directions, for L = |x - target|^2.
    x = res.x.cpu().numpy()
    target = 0.9 * x  # a coherent residual x - target, so the first-order change dominates the FD error
    gL = solver.equilibrium_vjp(dev, b.fixed, res.x, torch.from_numpy(2 * (x - target)).cuda(), opts).cpu().numpy()
    tight = solver.NewtonOptions(gradient_tolerance=1e-10)
    # Smooth, coherent directions (a random-sign field over 663,552 tets would leave a first-order change of order
    # sqrt(T) against third-order terms of order T h^3, i.e. an FD error far above h^2).
    centroid = b.rest[b.tets].mean(axis=1)
    fields = [rng.uniform(0.5, 1.5, size=(b.tets.shape[0], 1)),
              (1.0 + 0.5 * np.sin(2 * np.pi * centroid[:, 0]) * np.cos(np.pi * centroid[:, 2]))[:, None]]
    for d, field in enumerate(fields):
        dth = field * b.theta * 1e-3
        vals = []
        for s in (1.0, -1.0):
            th = torch.from_numpy(b.theta + s * dth).cuda()
            rs = solver.newton_solve(dev, b.fixed, res.x, tight, linear=lin, theta=th)
            assert rs.status == solver.NewtonStatus.Converged
            vals.append(float(((rs.x.cpu().numpy() - target) ** 2).sum()))
        fd = (vals[0] - vals[1]) / 2.0
        ad = float((gL * dth).sum())
        print(f"direction {d}: FD {fd!r} adjoint {ad!r}", flush=True)
        # Re-solves stop at |g| <= 1e-10 (a change of L far below the difference); the O(h^2) truncation of
        # coherent steps of 1e-3 theta is ~1e-6 relative.
        assert abs(fd - ad) <= 1e-4 * abs(ad)


def test_eight_beams_union_equals_separate_solves(torch):
    beams = workloads.beam_batch(8, seed=0)
    parts = [(b.num_vertices, b.tets, b.rest, b.theta) for b in beams]
    batch = BatchedFemEnergy(parts)
    fixed = np.concatenate([b.fixed for b in beams])
    lin = torch.from_numpy(np.concatenate([-b.load for b in beams])).cuda()
    x0 = torch.zeros(batch.num_dofs, dtype=torch.float64, device="cuda")
    res = solver.newton_solve(batch, fixed, x0, _opts(), linear=lin)
    print(f"union of 8 beams ({batch.num_tets} tets): status {res.status} newton {res.iterations} "
          f"|g| {res.gradient_norm:.3g}", flush=True)
    assert res.status == solver.NewtonStatus.Converged and res.gradient_norm < 1e-8
    x = res.x.cpu().numpy()
    off = 0
    for k, b in enumerate(beams):
        n = 3 * b.num_vertices
        dev = VolumetricFemEnergy(b.num_vertices, b.tets, b.rest, b.theta)
        one = solver.newton_solve(dev, b.fixed, torch.zeros(n, dtype=torch.float64, device="cuda"), _opts(),
                                  linear=torch.from_numpy(-b.load).cuda())
        assert one.status == solver.NewtonStatus.Converged
        xo = one.x.cpu().numpy()
        err = np.abs(x[off:off + n] - xo).max() / np.abs(xo).max()
        print(f"beam {k} {b.name}: union vs separate {err:.3g}", flush=True)
        # Both stop at |g|_inf < 1e-8 against loads of ~1e-2 per vertex on stiffness ~1e4: x agrees to ~1e-6.
        assert err <= 1e-5
        off += n
