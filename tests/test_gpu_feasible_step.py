"""max_feasible_step on the MI355X (through the C ABI) against the oracle and the reference's known answers.

Bar: within 1e-12 relative of the oracle (the same bisection on polynomial coefficients that agree to rounding);
exactly 1 where nothing limits the step; every state x + t dx, t in [0, alpha], evaluates without a domain error.
"""
from pathlib import Path

import numpy as np
import pytest

from litepgo_amd import BatchedFemEnergy, VolumetricFemEnergy, kuhn_grid, lame_parameters, workloads
from litepgo_amd import _native

pytestmark = pytest.mark.gpu

TOL = 1e-12
GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def test_two_tet_known_answers(torch, oracle_mod):
    d = np.load(GOLD / "two_tet.npz")
    X = d["X"].reshape(-1)
    dev = VolumetricFemEnergy(5, d["tets"], d["X"], d["theta"])
    zero = np.zeros(15)
    assert dev.max_feasible_step(zero, zero) == 1
    assert dev.max_feasible_step(zero, 0.1 * X) == 1
    direction = zero.copy()
    direction[0::3] = -4 * X[0::3]
    direction[1::3] = -4 * X[1::3]
    alpha = dev.max_feasible_step(zero, direction)
    assert abs(alpha - 0.225) <= 1e-6
    for i in range(21):
        assert np.isfinite(dev.eval_value((alpha * i / 20) * direction))
    direction[:] = 0
    direction[14] = 16
    assert abs(dev.max_feasible_step(zero, direction) - 0.1125) <= 1e-12
    t = torch.from_numpy(direction).cuda()
    assert abs(dev.max_feasible_step_device(torch.zeros(15, dtype=torch.float64, device="cuda"), t) - 0.1125) <= 1e-12
    with pytest.raises(ValueError):  # infeasible start: domain_error
        dev.max_feasible_step(-X, zero)
    with pytest.raises(ValueError):  # non-finite direction: invalid_argument
        dev.max_feasible_step(zero, np.full(15, np.nan))
    assert dev.max_feasible_step(zero, zero) == 1  # the handle recovers after errors


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_grid_random_directions_match_oracle(torch, oracle_mod, seed):
    X, tets = kuhn_grid(10)
    lam, mu = lame_parameters(1e5, 0.4)
    theta = np.tile([lam, mu], (tets.shape[0], 1))
    dev = VolumetricFemEnergy(len(X), tets, X, theta)
    ref = oracle_mod.OracleEnergy(len(X), tets, X, theta)
    rng = np.random.default_rng(seed)
    u = rng.uniform(-0.005, 0.005, size=X.size)
    for scale in (0.001, 0.05, 0.3, 3.0):
        dx = rng.normal(scale=scale, size=X.size)
        a = dev.max_feasible_step(u, dx)
        r = ref.max_feasible_step(u, dx)
        print(f"seed {seed} scale {scale}: device {a!r} oracle {r!r}", flush=True)
        if r == 1.0:
            assert a == 1.0
        else:
            assert abs(a - r) <= TOL * r
            assert np.isfinite(dev.eval_value(u + a * dx))


def test_equal_steps_everywhere_match_oracle(torch, oracle_mod):
    """A uniform compression along x of a 12^3 grid gives every tet the same step: thousands of lanes tie for the
    minimum, so the kernel's pruning (lanes whose step cannot reach the segment's least published bound stop
    bisecting) must keep a lane that attains it.  F = diag(1 - c t, 1, 1), so det F has the simple root t = 1 / c
    and the step is 0.9 / c up to the reference's 64 eps tolerance; the result equals the oracle's."""
    X, tets = kuhn_grid(12)
    lam, mu = lame_parameters(1e5, 0.4)
    theta = np.tile([lam, mu], (tets.shape[0], 1))
    dev = VolumetricFemEnergy(len(X), tets, X, theta)
    ref = oracle_mod.OracleEnergy(len(X), tets, X, theta)
    zero = np.zeros(X.size)
    for c in (1.25, 2.0, 4.0):
        dx = np.zeros_like(X)
        dx[:, 0] = -c * X[:, 0]
        dx = dx.reshape(-1)
        a = dev.max_feasible_step(zero, dx)
        r = ref.max_feasible_step(zero, dx)
        print(f"compression {c}: device {a!r} oracle {r!r}", flush=True)
        assert abs(a - r) <= TOL * r
        assert abs(a - 0.9 / c) <= 1e-12 / c


def test_full_size_step_property(torch):
    w = workloads.plain_large()
    dev = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
    rng = np.random.default_rng(5)
    x = torch.from_numpy(w.u).cuda()
    dx = torch.from_numpy(rng.normal(scale=0.05, size=w.u.size)).cuda()
    alpha = dev.max_feasible_step_device(x, dx)
    print(f"64^3 alpha {alpha!r}", flush=True)
    assert 0.0 < alpha < 1.0
    for t in (0.25 * alpha, 0.5 * alpha, alpha):
        dev.eval_device(x + t * dx, False, None)  # raises on det F <= 0
    # Run-to-run identical (a minimum is order-free).
    assert dev.max_feasible_step_device(x, dx) == alpha


def test_batch_per_mesh_steps(torch):
    meshes = workloads.batched(12, seed=3)
    batch = BatchedFemEnergy(meshes)
    rng = np.random.default_rng(7)
    xs = [w.u for w in meshes]
    dxs = [rng.normal(scale=0.05 * (1 + m % 3), size=w.u.size) for m, w in enumerate(meshes)]
    x = torch.from_numpy(np.concatenate(xs)).cuda()
    dx = torch.from_numpy(np.concatenate(dxs)).cuda()
    alphas = batch.max_feasible_step_per_mesh(x, dx).cpu().numpy()
    for m, w in enumerate(meshes):
        plain = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
        assert alphas[m] == plain.max_feasible_step(xs[m], dxs[m])
    assert batch.max_feasible_step_device(x, dx) == alphas.min()


def test_unused_vertex_is_checked(torch):
    X, tets = kuhn_grid(2)
    X = np.vstack([X, [[5.0, 5.0, 5.0]]])  # a vertex no tet uses
    lam, mu = lame_parameters(1e5, 0.4)
    dev = VolumetricFemEnergy(len(X), tets, X, np.tile([lam, mu], (tets.shape[0], 1)))
    x = np.zeros(X.size)
    dx = np.zeros(X.size)
    dx[-1] = np.inf
    with pytest.raises(ValueError):
        dev.max_feasible_step(x, dx)
    xt = torch.zeros(X.size, dtype=torch.float64, device="cuda")
    dxt = torch.from_numpy(dx).cuda()
    with pytest.raises(ValueError):
        dev.max_feasible_step_device(xt, dxt)
