"""The allocation-free evaluation (lpg_tet_energy_eval_device_noalloc, through the C ABI): on two calls in a row,
the second with a smaller mesh, its energy, gradient and Hessian values equal lpg_tet_energy_eval_device's bit for
bit for every Hessian kind; a batch handle reports the total of its meshes; a handle whose mesh is not on the
device yet and an evaluation that lists more tets for the eigen-analysis than the static scratch holds are
rejected with clear errors, after which both entry points still work."""
import ctypes

import numpy as np
import pytest

from litepgo_amd import BatchedFemEnergy, VolumetricFemEnergy, _native, workloads

pytestmark = pytest.mark.gpu
KINDS = {"none": -1, "exact": 0, "projected": 1}


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available()
    return torch


def noalloc(torch, h, x, kind):
    """One call of the allocation-free entry point; outputs start at -7 so unwritten entries show."""
    code = KINDS[kind]
    g = torch.full((h.num_dofs,), -7.0, dtype=torch.float64, device="cuda")
    H = torch.full((h.nnz,), -7.0, dtype=torch.float64, device="cuda") if code >= 0 else None
    value = np.full(1, np.nan)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    st = _native.lib().lpg_tet_energy_eval_device_noalloc(
        h._h, ctypes.c_void_p(x.data_ptr()), None, 1, code, ctypes.c_void_p(value.ctypes.data),
        ctypes.c_void_p(g.data_ptr()), None if H is None else ctypes.c_void_p(H.data_ptr()), stream)
    _native.check(st, "eval_device_noalloc")
    return float(value[0]), g, H


def handle(w):
    return VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)


@pytest.mark.parametrize("kind", ["none", "exact", "projected"])
def test_two_calls_match_the_allocating_entry_point(torch, kind):
    big = workloads.plain(20, 1e5, 0.45, 0.05, seed=3, name="noalloc_20^3")
    small = workloads.plain(6, 3e4, 0.3, 0.08, seed=4, name="noalloc_6^3")
    for w in (big, small):  # the second call is smaller: stale scratch of the first must not leak into it
        print(f"{kind} {w.name}", flush=True)
        h = handle(w)
        x = torch.from_numpy(w.u).cuda()
        v0, g0, H0 = h.eval_device(x, True, None if kind == "none" else kind)
        listed = h.last_indefinite_count
        v, g, H = noalloc(torch, h, x, kind)
        torch.cuda.synchronize()
        assert v == v0
        assert torch.equal(g, g0)
        if H0 is not None:
            assert torch.equal(H, H0)
        if kind == "projected":
            assert h.last_indefinite_count == listed > 0


def test_batch_handle_reports_the_total(torch):
    meshes = workloads.batched(num_meshes=6, seed=5, lo=2, hi=6)
    batch = BatchedFemEnergy(meshes)
    x = torch.from_numpy(np.concatenate([w.u for w in meshes])).cuda()
    v0, g0, H0 = VolumetricFemEnergy.eval_device(batch, x, True, "projected")
    v, g, H = noalloc(torch, batch, x, "projected")
    assert v == v0 and torch.equal(g, g0) and torch.equal(H, H0)


def test_rejections_leave_both_entry_points_working(torch):
    w = workloads.plain(8, 1e5, 0.4, 0.05, seed=6, name="noalloc_8^3")
    h = handle(w)
    x = torch.from_numpy(w.u).cuda()
    with pytest.raises(RuntimeError, match="not on the device yet"):
        noalloc(torch, h, x, "exact")
    v0, g0, H0 = h.eval_device(x, True, "exact")
    v, g, H = noalloc(torch, h, x, "exact")
    assert v == v0 and torch.equal(g, g0) and torch.equal(H, H0)
    # 64^3 under strong compression noise: far more tets need the eigen-analysis than the 64 MiB scratch holds.
    big = workloads.plain(64, 1e5, 0.48, 0.05, seed=7, name="noalloc_64^3")
    hb = handle(big)
    xb = torch.from_numpy(big.u).cuda()
    hb.eval_device(xb, True, "projected")
    print("listed", hb.last_indefinite_count, flush=True)
    with pytest.raises(ValueError, match="tets need the eigen-analysis, the scratch holds"):
        noalloc(torch, hb, xb, "projected")
    vb0, gb0, _ = hb.eval_device(xb, True, "exact")
    vb, gb, _ = noalloc(torch, hb, xb, "exact")
    assert vb == vb0 and torch.equal(gb, gb0)
    v, g, H = noalloc(torch, h, x, "projected")
    v0, g0, H0 = h.eval_device(x, True, "projected")
    assert v == v0 and torch.equal(g, g0) and torch.equal(H, H0)
