"""Device equilibrium solver (include/litepgo_fem_solver.h) on the MI355X against the oracle's restatement of
newton.cpp and static_equilibrium_response.cpp (oracle/newton.py, a direct sparse solver), on small cantilevers:
the equilibrium, the adjoint (VJP) and tangent (JVP) sensitivities, their transpose identity, finite differences
of re-solved equilibria, a batch of beams solved as one union, and the error statuses."""
import numpy as np
import pytest

from litepgo_amd import BatchedFemEnergy, VolumetricFemEnergy, workloads
from litepgo_amd import solver

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available()
    return torch


def opts(**kw):
    return solver.NewtonOptions(gradient_tolerance=TOL, **kw)


@pytest.fixture(scope="module")
def beam_case(torch, oracle_mod):
    from oracle import newton

    b = workloads.beam((6, 3, 3), seed=5)
    dev = VolumetricFemEnergy(b.num_vertices, b.tets, b.rest, b.theta)
    ref = oracle_mod.OracleEnergy(b.num_vertices, b.tets, b.rest, b.theta)
    n = 3 * b.num_vertices
    x0 = torch.zeros(n, dtype=torch.float64, device="cuda")
    lin = torch.from_numpy(-b.load).cuda()
    res = solver.newton_solve(dev, b.fixed, x0, opts(), linear=lin)
    print(f"device newton: status {res.status} iterations {res.iterations} pcg {res.pcg_iterations} "
          f"|g| {res.gradient_norm:.3g}", flush=True)
    st, xr, er, gr, itr = newton.newton_solve(ref, b.fixed, np.zeros(n), TOL, b=-b.load)
    print(f"oracle newton: status {st} iterations {itr} |g| {gr:.3g}", flush=True)
    return b, dev, ref, res, xr, lin


def test_equilibrium_matches_oracle(beam_case):
    b, dev, ref, res, xr, _ = beam_case
    assert res.status == solver.NewtonStatus.Converged and res.gradient_norm <= TOL
    x = res.x.cpu().numpy()
    assert np.all(x[b.fixed] == 0.0)
    assert np.abs(x - xr).max() <= 1e-7 * np.abs(xr).max()
    # The device energy at the device solution equals the oracle's energy at it.
    e_ref = ref.eval(x)[0] - b.load @ x
    assert abs(res.energy - e_ref) <= 1e-12 * abs(e_ref)


def test_solve_is_reproducible_bitwise(torch, beam_case):
    """Every reduction of the device solver has a fixed order: a second solve repeats the first bit for bit."""
    b, dev, _, res, _, lin = beam_case
    x0 = torch.zeros(3 * b.num_vertices, dtype=torch.float64, device="cuda")
    again = solver.newton_solve(dev, b.fixed, x0, opts(), linear=lin)
    assert again.iterations == res.iterations and again.pcg_iterations == res.pcg_iterations
    assert torch.equal(again.x, res.x) and again.energy == res.energy


def test_adjoint_and_tangent_match_oracle(torch, beam_case):
    from oracle import newton

    b, dev, ref, res, _, _ = beam_case
    x = res.x.cpu().numpy()
    rng = np.random.default_rng(2)
    w = rng.normal(size=x.size)
    g_dev = solver.equilibrium_vjp(dev, b.fixed, res.x, torch.from_numpy(w).cuda(), opts()).cpu().numpy()
    g_ref = newton.equilibrium_vjp(ref, x, b.fixed, w)
    assert np.abs(g_dev - g_ref).max() <= 1e-7 * np.abs(g_ref).max()
    v = rng.normal(size=b.theta.shape) * b.theta
    dx_dev = solver.equilibrium_jvp(dev, b.fixed, res.x, torch.from_numpy(v).cuda(), opts()).cpu().numpy()
    dx_ref = newton.equilibrium_jvp(ref, x, b.fixed, v)
    assert np.abs(dx_dev - dx_ref).max() <= 1e-7 * np.abs(dx_ref).max()
    # Transpose identity w . (J v) = (J^T w) . v on the device results.
    assert abs(w @ dx_dev - (g_dev * v).sum()) <= 1e-8 * abs(w @ dx_dev)


def test_adjoint_matches_finite_differences_of_device_solves(torch, beam_case):
    b, dev, _, res, _, lin = beam_case
    rng = np.random.default_rng(3)
    target = res.x.cpu().numpy() * (1 + 0.1 * rng.normal(size=res.x.numel()))
    r = res.x.cpu().numpy() - target
    # L = |x - target|^2 : dL/dx = 2 (x - target)
    g = solver.equilibrium_vjp(dev, b.fixed, res.x, torch.from_numpy(2 * r).cuda(), opts()).cpu().numpy()
    for _ in range(2):
        dth = rng.normal(size=b.theta.shape) * b.theta * 1e-3
        vals = []
        for s in (1.0, -1.0):
            th = torch.from_numpy(b.theta + s * dth).cuda()
            rs = solver.newton_solve(dev, b.fixed, res.x, opts(), linear=lin, theta=th)
            assert rs.status == solver.NewtonStatus.Converged
            xs = rs.x.cpu().numpy()
            vals.append(((xs - target) ** 2).sum())
        fd = (vals[0] - vals[1]) / 2.0
        ad = float((g * dth).sum())
        print(f"FD {fd!r} adjoint {ad!r}", flush=True)
        # Error budget: each re-solve stops at |g| <= 1e-10, which moves L by ~1e-10 against a difference of
        # ~1e-5 (steps of 1e-3 theta); the O(h^2) truncation is ~1e-6 relative.
        assert abs(fd - ad) <= 1e-4 * abs(ad)


def test_batch_of_beams_solved_as_one_union(torch):
    beams = [workloads.beam(dims, seed=s) for s, dims in enumerate([(4, 2, 2), (5, 3, 2), (3, 3, 3)])]
    parts = [(b.num_vertices, b.tets, b.rest, b.theta) for b in beams]
    batch = BatchedFemEnergy(parts)
    fixed = np.concatenate([b.fixed for b in beams])
    lin = torch.from_numpy(np.concatenate([-b.load for b in beams])).cuda()
    x0 = torch.zeros(batch.num_dofs, dtype=torch.float64, device="cuda")
    res = solver.newton_solve(batch, fixed, x0, opts(), linear=lin)
    assert res.status == solver.NewtonStatus.Converged
    x = res.x.cpu().numpy()
    off = 0
    for b in beams:
        n = 3 * b.num_vertices
        dev = VolumetricFemEnergy(*[b.num_vertices, b.tets, b.rest, b.theta])
        one = solver.newton_solve(dev, b.fixed, torch.zeros(n, dtype=torch.float64, device="cuda"), opts(),
                                  linear=torch.from_numpy(-b.load).cuda())
        xo = one.x.cpu().numpy()
        assert np.abs(x[off:off + n] - xo).max() <= 1e-7 * np.abs(xo).max()
        off += n


def test_statuses_and_errors(torch, beam_case):
    b, dev, _, res, _, lin = beam_case
    n = dev.num_dofs
    x0 = torch.zeros(n, dtype=torch.float64, device="cuda")
    r0 = solver.newton_solve(dev, b.fixed, x0, opts(max_iterations=0), linear=lin)
    assert r0.status == solver.NewtonStatus.MaxIterations and r0.iterations == 0
    with pytest.raises(ValueError):
        solver.newton_solve(dev, b.fixed, x0, solver.NewtonOptions(), linear=lin)  # tolerance not set
    with pytest.raises(ValueError):
        solver.newton_solve(dev, b.fixed, x0, opts(line_search_shrink=1.5), linear=lin)
    bad = torch.from_numpy(-np.repeat(b.rest, 1, axis=0).reshape(-1) * 2).cuda()  # x = -2 X: inverted
    with pytest.raises(ValueError):
        solver.newton_solve(dev, b.fixed, bad, opts(), linear=lin)
    again = solver.newton_solve(dev, b.fixed, res.x, opts(), linear=lin)  # the handle recovers
    assert again.status == solver.NewtonStatus.Converged and again.iterations == 0
