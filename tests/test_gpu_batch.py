"""Multi-mesh batch on the MI355X: every mesh's energy, gradient and Hessian values equal its own plain
evaluation bitwise (same launch shape) and the oracle to 1e-12; hipGraph replays match the eager result on every
replay with eager work in between; the autotuner changes no bit and caches its choice; errors of one mesh surface."""
import numpy as np
import pytest

from litepgo_amd import BatchedFemEnergy, VolumetricFemEnergy, workloads
from litepgo_amd._native import DomainError

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


@pytest.fixture(scope="module")
def meshes():
    return workloads.batched(num_meshes=24, seed=11, lo=2, hi=7)


def union_x(torch, meshes):
    return torch.from_numpy(np.concatenate([w.u for w in meshes])).cuda()


@pytest.mark.parametrize("kind", ["exact", "projected"])
def test_batch_equals_plain_per_mesh(torch, meshes, oracle_mod, kind):
    batch = BatchedFemEnergy(meshes)
    x = union_x(torch, meshes)
    E, g, H = batch.eval_device(x, True, kind)
    E, g, H = E.cpu().numpy(), g.cpu().numpy(), H.cpu().numpy()
    vo, _, no = batch.offsets()
    for m, w in enumerate(meshes):
        print(f"mesh {m} {w.name}", flush=True)
        plain = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
        v, gp, Hp = plain.eval_values(w.u, True, kind)
        assert E[m] == v, (m, E[m], v)
        assert np.array_equal(g[3 * vo[m]:3 * vo[m + 1]], gp)
        assert np.array_equal(H[no[m]:no[m + 1]], Hp)
        if m < 4:
            ref = oracle_mod.OracleEnergy(w.num_vertices, w.tets, w.rest, w.theta)
            vr, gr, Hr = ref.eval(w.u, True, kind)
            assert abs(v - vr) <= TOL * abs(vr) and rel(gp, gr) <= TOL and rel(Hp, Hr) <= TOL


def test_union_through_single_energy_entry_points(torch, meshes):
    """The batch handle used as one energy (VolumetricFemEnergy methods, host and device) reports the total of its
    meshes' energies, and its status word stays intact (24 segment energies no longer land in the scalar slot)."""
    batch = BatchedFemEnergy(meshes)
    x = union_x(torch, meshes)
    E, _, _ = batch.eval_device(x, True, "projected")
    per_mesh = E.cpu().numpy()
    total, g, _ = VolumetricFemEnergy.eval_device(batch, x, True, "projected")
    assert abs(total - per_mesh.sum()) <= TOL * abs(per_mesh).sum()
    v, gh, _ = batch.eval_values(x.cpu().numpy(), True, "exact")
    assert abs(v - per_mesh.sum()) <= 1e-10 * abs(per_mesh).sum()
    for _ in range(3):  # repeated calls: no stale status from overwritten scalars
        total2, _, _ = VolumetricFemEnergy.eval_device(batch, x, False, None)
        assert total2 == total or abs(total2 - total) <= TOL * abs(total)


def test_graph_replays_match_eager(torch, meshes):
    batch = BatchedFemEnergy(meshes)
    x = union_x(torch, meshes)
    E0, g0, H0 = (t.clone() for t in batch.eval_device(x, True, "projected"))
    batch.set_graph(True)
    E = torch.empty_like(E0)
    g = torch.empty_like(g0)
    H = torch.empty_like(H0)
    junk = torch.zeros(1 << 20, dtype=torch.float64, device="cuda")
    for r in range(5):
        print(f"replay {r}", flush=True)
        E.fill_(-1.0)
        g.fill_(-1.0)
        H.fill_(-1.0)
        batch.eval_device(x, True, "projected", out_E=E, out_g=g, out_H=H)
        torch.cuda.synchronize()
        assert torch.equal(E, E0) and torch.equal(g, g0) and torch.equal(H, H0), r
        junk.add_(1.0)  # eager work between replays
        torch.cuda.synchronize()
    # A different output buffer falls back to the eager path and is still right.
    E2, g2, H2 = batch.eval_device(x, True, "projected")
    assert torch.equal(E2, E0) and torch.equal(g2, g0) and torch.equal(H2, H0)


def test_graph_replays_report_status_and_recover(torch, meshes):
    """In graph mode a batch's replay hands energy and status to the host from a read-back kernel: a non-finite x
    raises the reference's error class from a replay, and the next replay with good contents is exact again."""
    import math

    batch = BatchedFemEnergy(meshes)
    x = union_x(torch, meshes)
    E0, g0, H0 = (t.clone() for t in batch.eval_device(x, True, "projected"))
    E = torch.empty_like(E0)
    g = torch.empty_like(g0)
    H = torch.empty_like(H0)
    xr = x.clone()
    batch.set_graph(True)
    try:
        for _ in range(2):  # eager, then captured
            batch.eval_device(xr, True, "projected", out_E=E, out_g=g, out_H=H)
        for r in range(3):
            xr.copy_(x)
            xr[5] = math.nan
            with pytest.raises(ValueError, match="finite"):
                batch.eval_device(xr, True, "projected", out_E=E, out_g=g, out_H=H)
            xr.copy_(x)
            batch.eval_device(xr, True, "projected", out_E=E, out_g=g, out_H=H)
            torch.cuda.synchronize()
            assert torch.equal(E, E0) and torch.equal(g, g0) and torch.equal(H, H0), r
    finally:
        batch.set_graph(False)


def test_autotune_keeps_results_and_caches_its_choice(torch, meshes):
    """The autotuner's choice changes no bit; a second handle with the same mesh-size histogram takes the cached
    choice without timing, and a negative repeat count times again."""
    batch = BatchedFemEnergy(meshes)
    x = union_x(torch, meshes)
    E0, g0, H0 = (t.cpu().numpy() for t in batch.eval_device(x, True, "projected"))
    shape, ms = batch.autotune(x, "projected", repeats=-2)
    print("chosen", shape, ms, flush=True)
    assert ms > 0 and batch.get_launch() == shape[:4]
    E, g, H = (t.cpu().numpy() for t in batch.eval_device(x, True, "projected"))
    assert np.array_equal(E, E0) and np.array_equal(g, g0) and np.array_equal(H, H0)
    other = BatchedFemEnergy(meshes)
    shape2, ms2 = other.autotune(x, "projected", repeats=2)
    assert shape2 == shape and ms2 == ms  # cached: same choice, same recorded time
    assert other.get_launch() == shape[:4]
    E2, g2, H2 = (t.cpu().numpy() for t in other.eval_device(x, True, "projected"))
    assert np.array_equal(E2, E0) and np.array_equal(g2, g0) and np.array_equal(H2, H0)


def test_error_in_one_mesh_fails_the_batch(torch, meshes):
    batch = BatchedFemEnergy(meshes)
    xs = [w.u.copy() for w in meshes]
    w = meshes[5]
    xs[5] = (-1.5 * w.rest).reshape(-1)  # inverts every tet of mesh 5
    x = torch.from_numpy(np.concatenate(xs)).cuda()
    with pytest.raises(DomainError):
        batch.eval_device(x, True, "exact")
    # the handle recovers
    batch.eval_device(union_x(torch, meshes), True, "exact")


def test_enqueued_per_mesh_evaluations_equal_synchronous_ones(torch, meshes):
    """Evaluations of many plain handles enqueued on one stream and waited for afterwards give the bits of the
    synchronous calls; the handle is busy in between; waiting twice fails; a batch handle reports per-mesh
    energies through the same pair."""
    from litepgo_amd._native import BusyError

    handles = [VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta) for w in meshes]
    xs = [torch.from_numpy(w.u.copy()).cuda() for w in meshes]
    ref = [h.eval_device(x, True, "projected") for h, x in zip(handles, xs)]
    outs = [h.eval_enqueue(x, True, "projected") for h, x in zip(handles, xs)]
    with pytest.raises(BusyError):
        handles[0].eval_device(xs[0], True, "projected")
    values = [h.eval_wait() for h in handles]
    for m, ((v, g, H), (_, g2, H2), v2) in enumerate(zip(ref, outs, values)):
        assert v2 == v and torch.equal(g2, g) and torch.equal(H2, H), m
    with pytest.raises(RuntimeError, match="no evaluation was enqueued"):
        handles[0].eval_wait()
    batch = BatchedFemEnergy(meshes)
    x = union_x(torch, meshes)
    E0, g0, H0 = batch.eval_device(x, True, "exact")
    E = torch.empty_like(E0)
    _, g, H = batch.eval_enqueue(x, True, "exact", out_E=E)
    assert batch.eval_wait() is None
    assert torch.equal(E, E0) and torch.equal(g, g0) and torch.equal(H, H0)
    batch.eval_enqueue(x, False, None)
    total = batch.eval_wait()  # no energies buffer: the sum over the meshes
    assert abs(total - float(E0.sum())) <= TOL * float(E0.abs().sum())
