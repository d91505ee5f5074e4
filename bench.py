#!/usr/bin/env python3
"""Benchmark of the neo-Hookean Tet P1 assembly on one MI355X.

One step = one Energy::eval(x, gradient=True, hessian=Projected) of the 64^3 Kuhn-grid workload (1,572,864 tets,
274,625 vertices, 36,168,777 Hessian values; E = 1e5, nu = 0.45, 20% stretch in x plus U(+-0.02 h) noise): energy,
gradient and the PSD-projected Hessian assembled into the CSR value array, with inputs and outputs resident in HBM.
`value` is tets/s of the whole step (every kernel, the status/energy read-back and the host synchronisation).

Also reported: the assembly kernel against the HBM roofline (HIP events around the kernel on its stream) and its
FP64 rate against the FP64 vector peak, the per-phase times, the same step with the Exact Hessian, the 256-mesh
batched evaluation (autotuned launch shape, replayed as a hipGraph) with its own assembly roofline, against a
per-mesh loop of synchronous calls and one of enqueued calls with a single wait, the 32^3 differentiation workload's backward time against its forward time, the equilibrium-sensitivity
problem (device Newton + Jacobi-PCG + adjoint on a 48^3 cantilever; 8 uneven beams with --equilibrium-batch), and
the oracle (C++ CPU restatement of the reference algorithm) timed on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batched]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "tets/s neo-Hookean E/grad/Hess+CSR assembly; batched & backward vs plain"
HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, vendor figure
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 vector, vendor figure
# Algorithmic (compulsory) bytes per assembly: 176 B per tet (connectivity 16, Dm^-1 72, w 8, (lambda, mu) 16,
# 16 int32 block slots 64), 48 B per vertex (x read, g write), 8 B per Hessian value written.
BYTES_PER_TET, BYTES_PER_VERTEX, BYTES_PER_NNZ = 176, 48, 8
PMC_SUMMARY = ROOT / "profiles" / "assemble_projected_64cube_pmc.txt"


def algorithmic_bytes(T, V, nnz):
    return BYTES_PER_TET * T + BYTES_PER_VERTEX * V + BYTES_PER_NNZ * nnz


def read_summary_value(path, key):
    """A `key: value` figure from a committed rocprofv3 summary, or None."""
    if not path.exists():
        return None
    m = re.search(rf"^{key}\s*[:=]\s*([0-9.eE+]+)", path.read_text(), re.M)
    return float(m.group(1)) if m else None


def cpu_model():
    """The host CPU's model name from /proc/cpuinfo (or platform.processor())."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


def _time_oracle(w, tets, nthreads, repeats):
    """Median wall time of `repeats` oracle evaluations E+g+H(Projected) of the workload restricted to `tets`, after
    one untimed warm-up evaluation."""
    from oracle import oracle

    ref = oracle.OracleEnergy(w.num_vertices, w.tets[tets], w.rest, w.theta[tets], nthreads=nthreads)
    ref.eval(w.u, True, "projected", nthreads=nthreads)
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        ref.eval(w.u, True, "projected", nthreads=nthreads)
        times.append(time.perf_counter() - t0)
    return statistics.median(times)


def cpu_baseline(w, nthreads, repeats, single_layers=2, n=64):
    """The oracle (oracle/liblpg_oracle.so, the C++ restatement of the reference's dense element route, 12x12
    eigen-clamp and column-owner assembly) timed on this host's cores, after a warm-up, median of `repeats`:
    at `nthreads` threads on the full 64^3 workload, and at 1 thread on a bounded sample of it (the tets of the
    first `single_layers` cell layers along x, same deformation)."""
    from oracle import oracle

    oracle.build()
    T = w.tets.shape[0]
    med = _time_oracle(w, np.arange(T), nthreads, repeats)
    T1 = single_layers * n * n * 6  # cells run x-slowest, six tets each
    med1 = _time_oracle(w, np.arange(T1), 1, repeats)
    affinity = len(os.sched_getaffinity(0))
    return {"value": T / med, "unit": "tets/s", "cores": nthreads, "kind": "port",
            "sample": f"oracle E+g+H(Projected)+CSR on the full 64^3 workload (1,572,864 tets), {nthreads} threads, "
                      f"median of {repeats} evals after a warm-up ({med:.2f} s per eval)",
            "single_thread": {"value": T1 / med1, "unit": "tets/s", "cores": 1,
                              "sample": f"the first {single_layers} x-layers of the same workload ({T1} tets), "
                                        f"median of {repeats} after a warm-up ({med1:.3f} s per eval)"},
            "host": {"cpu_model": cpu_model(), "nproc": os.cpu_count(), "affinity_cores": affinity,
                     "threads_used": nthreads,
                     "note": "threads capped at 16, the host-CPU share of one GPU on this machine"},
            "published_reference": {"value": 0.834e6, "unit": "tets/s",
                                    "what": "reference Exact Hessian (CSC), 10,368 tets, Apple M1 Pro, 1 thread "
                                            "(docs/public/benchmark-data/torch-material-baseline/timings.csv:24)"}}


def time_steps(dev, step, kind, steps, warmup):
    """Wall time per step (the evaluation replayed as one captured hipGraph, no timing events), then the per-phase
    kernel times from HIP events over a second run of the same steps (timing mode runs the launches eagerly)."""
    import torch

    dev.set_graph(True)
    for _ in range(max(warmup, 2)):  # the second call captures the graph
        step(kind)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(kind)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    dev.set_graph(False)
    dev.set_timing(True)
    for _ in range(steps):
        step(kind)
    torch.cuda.synchronize()
    ms, evals = dev.kernel_times()
    dev.set_timing(False)
    return elapsed / steps * 1e3, [m / evals for m in ms]


def bench_batched(steps, warmup):
    """256 uneven meshes (4..24 cells per axis), E+g+H(Projected): one batched evaluation (autotuned launch shape,
    captured hipGraph) against loops that evaluate each mesh with its own handle, synchronously and enqueued on
    one stream with the waits after the last launch."""
    import numpy as np
    import torch

    from litepgo_amd import BatchedFemEnergy, VolumetricFemEnergy, workloads

    meshes = workloads.batched(256)
    T = sum(w.tets.shape[0] for w in meshes)
    batch = BatchedFemEnergy(meshes)
    x = torch.from_numpy(np.concatenate([w.u for w in meshes])).cuda()
    shape, tuned_ms = batch.autotune(x, "projected", repeats=3)
    batch.set_graph(True)
    E = torch.empty(batch.num_meshes, dtype=torch.float64, device="cuda")
    g = torch.empty(batch.num_dofs, dtype=torch.float64, device="cuda")
    H = torch.empty(batch.nnz, dtype=torch.float64, device="cuda")
    for _ in range(warmup):
        batch.eval_device(x, True, "projected", out_E=E, out_g=g, out_H=H)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.eval_device(x, True, "projected", out_E=E, out_g=g, out_H=H)
    torch.cuda.synchronize()
    batch_ms = (time.perf_counter() - t0) / steps * 1e3
    # Per-phase kernel times of the same evaluation (eager launches with HIP events on the handle's stream).
    batch.set_graph(False)
    batch.set_timing(True)
    for _ in range(steps):
        batch.eval_device(x, True, "projected", out_E=E, out_g=g, out_H=H)
    torch.cuda.synchronize()
    bms, bevals = batch.kernel_times()
    batch.set_timing(False)
    bphase = [m / bevals for m in bms]
    B = algorithmic_bytes(T, batch.num_vertices, batch.nnz)  # the sum of the meshes' own algorithmic bytes

    plains = [VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta) for w in meshes]
    xs = [torch.from_numpy(w.u).cuda() for w in meshes]
    gs = [torch.empty(p.num_dofs, dtype=torch.float64, device="cuda") for p in plains]
    Hs = [torch.empty(p.nnz, dtype=torch.float64, device="cuda") for p in plains]

    def loop():
        for p, xm, gm, Hm in zip(plains, xs, gs, Hs):
            p.eval_device(xm, True, "projected", out_g=gm, out_H=Hm)

    def async_loop():
        # Every mesh's evaluation enqueued on the current stream, then one wait per handle (the first wait
        # synchronises the stream; the others read back what is already there).
        for p, xm, gm, Hm in zip(plains, xs, gs, Hs):
            p.eval_enqueue(xm, True, "projected", out_g=gm, out_H=Hm)
        for p in plains:
            p.eval_wait()

    def timed(fn, n):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n * 1e3

    loop_steps = max(2, steps // 4)
    loop_ms = timed(loop, loop_steps)
    async_ms = timed(async_loop, loop_steps)
    return {"meshes": len(meshes), "tets": T, "vertices": batch.num_vertices, "hessian_nnz": batch.nnz,
            "ms_per_step": batch_ms, "tets_per_s": T / (batch_ms * 1e-3),
            "phases_ms": {"psd_test_and_eigen_clamp": bphase[0], "assembly": bphase[1],
                          "energy_reduction": bphase[2], "all_kernels": bphase[3]},
            "roofline": {"bound": "hbm", "kernel": ("k_assemble_persistent" if shape[4] else "k_assemble") +
                                               " (Projected, gradient), all meshes",
                         "achieved": B / (bphase[1] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": B / (bphase[1] * 1e-3) / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes": B,
                         "kernel_ms": bphase[1]},
            "per_mesh_loop_ms": loop_ms, "per_mesh_loop_tets_per_s": T / (loop_ms * 1e-3),
            "speedup_vs_per_mesh_loop": loop_ms / batch_ms,
            "per_mesh_async_loop_ms": async_ms, "per_mesh_async_loop_tets_per_s": T / (async_ms * 1e-3),
            "speedup_vs_per_mesh_async_loop": async_ms / batch_ms,
            "launch": {"lanes_per_vertex": shape[0], "waves": shape[1], "vertices_per_tile": shape[2],
                       "lds_budget": shape[3], "persistent_grid": shape[4], "graph": True,
                       "autotune_ms": tuned_ms}}


def bench_differentiation(steps, warmup):
    """32^3 differentiation workload (per-tet log-normal E, uniform nu).  Forward: theta = lame(E, nu), (E, f) by
    one assembly, and the objective L = E + w . f (w fixed, random).  Backward: dL/dx, dL/dE_young, dL/dnu by
    autograd (one fused library call for s f + H w and s dE/dtheta + (df/dtheta)^T w, plus the (lambda, mu)
    chain rule).  Wall-clock medians, the host synchronised with the device on both sides of each part."""
    import numpy as np
    import torch

    from litepgo_amd import VolumetricFemEnergy, workloads
    from litepgo_amd import torch as lt

    w, E_young, nu = workloads.differentiate(seed=0, n=32)
    pot = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
    x = torch.from_numpy(w.u).cuda().requires_grad_(True)
    Et = torch.from_numpy(E_young).cuda().requires_grad_(True)
    nut = torch.from_numpy(nu).cuda().requires_grad_(True)
    wt = torch.from_numpy(np.random.default_rng(1).normal(size=pot.num_dofs)).cuda()

    def forward():
        E, f = lt.energy_and_force(pot, x, lt.lame(Et, nut))
        return E + torch.dot(wt, f)

    def backward(objective):
        return torch.autograd.grad(objective, (x, Et, nut))

    for _ in range(warmup):
        backward(forward())
    fwd, bwd = [], []
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        objective = forward()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        backward(objective)
        torch.cuda.synchronize()
        fwd.append(t1 - t0)
        bwd.append(time.perf_counter() - t1)
    f_ms, b_ms = statistics.median(fwd) * 1e3, statistics.median(bwd) * 1e3
    T = pot.num_tets
    return {"workload": "32^3 Kuhn grid, per-tet E log-normal(ln 1e5, 0.5), nu U[0.3, 0.45], noise +-0.05h; "
                        "objective E + w.f, gradients in x, E_young, nu",
            "tets": T, "forward_ms": f_ms, "backward_ms": b_ms, "backward_over_forward": b_ms / f_ms,
            "forward_tets_per_s": T / (f_ms * 1e-3), "backward_tets_per_s": T / (b_ms * 1e-3)}


def bench_side_products(dev, w, steps, warmup):
    """The operator's other entry points on the headline workload, wall-clock medians with the host synchronised on
    both sides (each call ends with its status read-back): the feasible step along a random direction, and the
    force response to per-call plastic parts (Fp = I + 0.05 N(0, 1) per tet): value, JVP and VJP."""
    import numpy as np
    import torch

    rng = np.random.default_rng(3)
    T = dev.num_tets
    x = torch.from_numpy(w.u).cuda()
    dx = torch.from_numpy(rng.normal(scale=0.05, size=dev.num_dofs)).cuda()
    Fp = np.eye(3)[None] + 0.05 * rng.normal(size=(T, 3, 3))
    fp = torch.from_numpy(np.ascontiguousarray(Fp.transpose(0, 2, 1).reshape(T, 9))).cuda()
    dfp = torch.from_numpy(rng.normal(size=(T, 9))).cuda()
    wv = torch.from_numpy(rng.normal(size=dev.num_dofs)).cuda()
    calls = {
        "feasible_step": lambda: dev.max_feasible_step_device(x, dx),
        "force_fp": lambda: dev.force_fp_device(x, fp),
        "force_jvp_fp": lambda: dev.force_jvp_fp_device(x, dfp, fp),
        "force_vjp_fp": lambda: dev.force_vjp_fp_device(x, wv, fp),
    }
    out = {"workload": "plain 64^3 Kuhn grid (as the headline); Fp = I + 0.05 N(0, 1) per tet"}
    for name, call in calls.items():
        for _ in range(warmup):
            call()
        times = []
        for _ in range(steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            call()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        ms = statistics.median(times) * 1e3
        out[name] = {"ms": ms, "tets_per_s": T / (ms * 1e-3)}
    return out


def bench_equilibrium(batch):
    """Equilibrium sensitivity (cantilever, x = 0 face fixed, gravity 9.8 along -z, density 1000, per-tet E
    log-normal(ln 1e5, 0.5), nu 0.45): device Newton with Jacobi-PCG to |g|_inf < 1e-8 from rest, the same for the
    target (a seeded perturbation of the E field), then the adjoint gradient of |x - x_target|^2 with respect to
    the per-tet E field.  batch=False: one 48^3 beam; True: 8 beams with per-axis counts in 40..48, solved as one
    union handle."""
    import numpy as np
    import torch

    from litepgo_amd import BatchedFemEnergy, VolumetricFemEnergy, solver, workloads
    from litepgo_amd import torch as lt

    beams = workloads.beam_batch(8, seed=0) if batch else [workloads.beam(48, seed=0)]
    parts = [(b.num_vertices, b.tets, b.rest, b.theta) for b in beams]
    T = sum(b.tets.shape[0] for b in beams)
    en = BatchedFemEnergy(parts) if batch else VolumetricFemEnergy(*parts[0])
    fixed = solver.fixed_mask(en.num_dofs, np.concatenate([b.fixed for b in beams]))
    lin = torch.from_numpy(np.concatenate([-b.load for b in beams])).cuda()
    nu = torch.from_numpy(np.concatenate([b.nu for b in beams])).cuda()
    E = torch.from_numpy(np.concatenate([b.E for b in beams])).cuda()
    theta = lt.lame(E, nu).detach()
    theta_t = lt.lame(torch.from_numpy(np.concatenate([b.E_target for b in beams])).cuda(), nu).detach()
    x0 = torch.zeros(en.num_dofs, dtype=torch.float64, device="cuda")
    opts = solver.NewtonOptions(gradient_tolerance=1e-8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = solver.newton_solve(en, fixed, x0, opts, linear=lin, theta=theta)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    res_t = solver.newton_solve(en, fixed, x0, opts, linear=lin, theta=theta_t)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    Er = E.clone().requires_grad_(True)
    g_theta = solver.equilibrium_vjp(en, fixed, res.x, 2.0 * (res.x - res_t.x), opts, theta=theta)
    (g_E,) = torch.autograd.grad(lt.lame(Er, nu), Er, grad_outputs=g_theta)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    ok = res.status == 0 and res_t.status == 0
    return {"workload": ("8 cantilevers, 40..48 cells per axis" if batch else "48^3 cantilever") +
                        ", x=0 fixed, gravity -9.8 z, density 1000, E log-normal(ln 1e5, 0.5), nu 0.45",
            "tets": T, "dofs": en.num_dofs, "converged": ok,
            "newton_iterations": res.iterations, "pcg_iterations": res.pcg_iterations,
            "gradient_norm": res.gradient_norm, "solve_s": t1 - t0, "target_solve_s": t2 - t1,
            "adjoint_s": t3 - t2, "pcg_iterations_per_s": res.pcg_iterations / (t1 - t0),
            "objective": float(((res.x - res_t.x) ** 2).sum()), "grad_E_norm": float(g_E.norm())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-repeats", type=int, default=9)
    ap.add_argument("--exact-only", action="store_true", help="time only the Exact-Hessian step (profiling aid)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-mesh batched comparison")
    ap.add_argument("--no-differentiation", action="store_true", help="skip the 32^3 forward/backward timing")
    ap.add_argument("--no-equilibrium", action="store_true", help="skip the 48^3 equilibrium-sensitivity solve")
    ap.add_argument("--no-side", action="store_true", help="skip the feasible-step and plastic-response timings")
    ap.add_argument("--equilibrium-batch", action="store_true",
                    help="also solve the 8-beam (40..48 cells per axis) equilibrium batch")
    ap.add_argument("--probe-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a known-size kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the operator runs on one device, "
              "and a multi-GPU run would give each rank its own meshes or its own share of one mesh.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from litepgo_amd import VolumetricFemEnergy, workloads

    assert torch.cuda.is_available(), "no HIP device"
    w = workloads.plain_large()
    T, V = w.tets.shape[0], w.num_vertices
    dev = VolumetricFemEnergy(V, w.tets, w.rest, w.theta)
    nnz = dev.nnz
    x = torch.from_numpy(w.u).cuda()
    g = torch.empty(dev.num_dofs, dtype=torch.float64, device="cuda")
    H = torch.empty(nnz, dtype=torch.float64, device="cuda")

    def step(kind):
        return dev.eval_device(x, True, kind, out_g=g, out_H=H)

    if args.probe_copy_mb:
        src = torch.ones(args.probe_copy_mb * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    B = algorithmic_bytes(T, V, nnz)
    e_ms, e_phase = time_steps(dev, step, "exact", args.steps, args.warmup)
    exact = {"tets_per_s": T / (e_ms * 1e-3), "ms_per_step": e_ms, "assemble_ms": e_phase[1],
             "assemble_hbm_frac": B / (e_phase[1] * 1e-3) / 1e9 / HBM_PEAK_GBS}
    if args.exact_only:
        print(json.dumps({"exact": exact}), flush=True)
        return
    p_ms, p_phase = time_steps(dev, step, "projected", args.steps, args.warmup)
    indefinite = dev.last_indefinite_count
    # The host-buffer entry point (x in, g and the Hessian values out through pageable host memory): PCIe-inclusive,
    # reported beside the headline, never as it.
    dev.eval_values(w.u, True, "projected")
    t0 = time.perf_counter()
    for _ in range(3):
        dev.eval_values(w.u, True, "projected")
    host_ms = (time.perf_counter() - t0) / 3 * 1e3
    assemble_ms = p_phase[1]
    achieved = B / (assemble_ms * 1e-3) / 1e9
    flops = read_summary_value(PMC_SUMMARY, "fp64_flops_per_launch")

    result = {
        "metric": METRIC,
        "value": T / (p_ms * 1e-3),
        "unit": "tets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": p_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain 64^3 Kuhn grid, E+g+H(PSD-projected)+CSR assembly, E=1e5 nu=0.45, "
                               "20% x-stretch + U(+-0.02h)", "tets": T, "vertices": V, "hessian_nnz": nnz,
                   "launch": list(dev.get_launch())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": read_summary_value(PMC_SUMMARY, "traffic_bytes_per_launch"),
                     "kernel": "k_assemble (Projected, gradient)", "kernel_ms": assemble_ms,
                     "algorithmic_bytes": B},
        "fp64": None if flops is None else {
            "achieved": flops / (assemble_ms * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": flops / (assemble_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, "flops_per_launch": flops},
        "phases_ms": {"psd_test_and_eigen_clamp": p_phase[0], "assembly": p_phase[1],
                      "energy_reduction": p_phase[2], "all_kernels": p_phase[3]},
        "indefinite_tets": indefinite, "indefinite_fraction": indefinite / T,
        "host_buffers": {"ms_per_step": host_ms, "tets_per_s": T / (host_ms * 1e-3),
                         "note": "lpg_tet_energy_eval with host x, g and H (PCIe-inclusive); not the headline"},
        "exact": exact,
    }
    if not args.no_side:
        result["side_products"] = bench_side_products(dev, w, args.steps, args.warmup)
    if not args.no_batched:
        result["batched"] = bench_batched(args.steps, args.warmup)
    if not args.no_differentiation:
        result["differentiation"] = bench_differentiation(args.steps, args.warmup)
    if not args.no_equilibrium:
        result["equilibrium"] = bench_equilibrium(batch=False)
        if args.equilibrium_batch:
            result["equilibrium_batch"] = bench_equilibrium(batch=True)
    if not args.no_cpu:
        nthreads = min(16, len(os.sched_getaffinity(0)))
        result["cpu_baseline"] = cpu_baseline(w, nthreads, args.cpu_repeats)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
