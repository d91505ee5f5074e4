"""One call at a time per energy object (src/core/common/busy_scope.hpp:11-26, mesh_energy.cpp:49; the reference's
tests/python/test_response_execution.py:10-46 checks the same guard): while one thread's call runs on a handle,
another thread's call on it fails with the logic error "MeshEnergy is already in use by another call"
(RuntimeError), the running call completes normally, and afterwards the handle evaluates exactly as before.
Different handles may be evaluated concurrently."""
import threading
import time

import numpy as np
import pytest

from litepgo_amd import VolumetricFemEnergy, solver, workloads
from litepgo_amd._native import BusyError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available()
    return torch


def _run_overlapped(long_call, probe):
    """Runs long_call in a worker thread (retrying if the probe got in first) and probes the handle from this thread
    until the worker ends; returns (worker result, busy errors seen by the probe, probe calls that succeeded)."""
    out, errs = {}, []

    def worker():
        for _ in range(50):
            try:
                out["result"] = long_call()
                return
            except BusyError:
                time.sleep(0.002)
            except Exception as e:  # noqa: BLE001 -- reported by the test
                errs.append(e)
                return

    t = threading.Thread(target=worker)
    t.start()
    busy, ok = [], 0
    while t.is_alive():
        try:
            probe()
            ok += 1
        except BusyError as e:
            busy.append(str(e))
        time.sleep(0.001)
    t.join()
    assert not errs, errs
    return out.get("result"), busy, ok


def test_overlapping_calls_on_one_handle_fail_busy_and_recover(torch):
    b = workloads.beam(24, seed=1)
    dev = VolumetricFemEnergy(b.num_vertices, b.tets, b.rest, b.theta)
    n = dev.num_dofs
    lin = torch.from_numpy(-b.load).cuda()
    x0 = torch.zeros(n, dtype=torch.float64, device="cuda")
    h = 1.0 / 24
    xs = torch.from_numpy(np.random.default_rng(0).uniform(-0.05 * h, 0.05 * h, size=n) * (~b.fixed)).cuda()
    E0, g0, H0 = dev.eval_device(xs, True, "projected")
    g0, H0 = g0.clone(), H0.clone()
    opts = solver.NewtonOptions(gradient_tolerance=1e-9)

    res, busy, ok = _run_overlapped(lambda: solver.newton_solve(dev, b.fixed, x0, opts, linear=lin),
                                    lambda: dev.eval_device(xs, True, "projected"))
    print(f"probe: {len(busy)} busy errors, {ok} calls completed", flush=True)
    assert res is not None and res.status == solver.NewtonStatus.Converged
    assert busy, "no call overlapped the running solve"
    assert all("MeshEnergy is already in use by another call" in m for m in busy), busy[:3]
    # The handle recovers: the same evaluation repeats the first one bit for bit.
    E1, g1, H1 = dev.eval_device(xs, True, "projected")
    assert E1 == E0 and torch.equal(g1, g0) and torch.equal(H1, H0)
    # Every entry point takes the guard: a derivative product overlapping the solve fails the same way.
    w = torch.from_numpy(np.random.default_rng(1).normal(size=n)).cuda()
    res, busy, _ = _run_overlapped(lambda: solver.newton_solve(dev, b.fixed, x0, opts, linear=lin),
                                   lambda: dev.hvp_device(xs, w))
    assert res is not None and busy


def test_different_handles_evaluate_concurrently(torch):
    w = workloads.plain(12, 1e5, 0.4, 0.05, seed=2)
    a = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
    b = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
    x = torch.from_numpy(w.u).cuda()
    ref = a.eval_device(x, True, "projected")
    ref = (ref[0], ref[1].clone(), ref[2].clone())
    errs, results = [], []

    def run(h):
        try:
            for _ in range(20):
                v, g, H = h.eval_device(x, True, "projected")
                results.append((v, g.clone(), H.clone()))
        except Exception as e:  # noqa: BLE001
            errs.append(e)

    ts = [threading.Thread(target=run, args=(h,)) for h in (a, b)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs, errs
    assert len(results) == 40
    for v, g, H in results:
        assert v == ref[0] and torch.equal(g, ref[1]) and torch.equal(H, ref[2])
