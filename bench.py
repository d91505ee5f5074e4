#!/usr/bin/env python3
"""Benchmark of the neo-Hookean Tet P1 assembly on one MI355X.

One step = one Energy::eval(x, gradient=True, hessian=Projected) of the 64^3 Kuhn-grid workload (1,572,864 tets,
274,625 vertices, 36,168,777 Hessian values; E = 1e5, nu = 0.45, 20% stretch in x plus U(+-0.02 h) noise): energy,
gradient and the PSD-projected Hessian assembled into the CSR value array, with inputs and outputs resident in HBM.
`value` is tets/s of the whole step (every kernel, the status/energy read-back and the host synchronisation).

Also reported: the assembly kernel against the HBM roofline (HIP events around the kernel on its stream) and its
FP64 rate against the FP64 vector peak, the per-phase times, the same step with the Exact Hessian, and the oracle
(C++ CPU restatement of the reference algorithm) timed on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
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


def cpu_baseline(w, nthreads, repeats):
    """The oracle (oracle/liblpg_oracle.so) on the full 64^3 workload, E+g+H(Projected), median of `repeats`."""
    from oracle import oracle

    oracle.build()
    ref = oracle.OracleEnergy(w.num_vertices, w.tets, w.rest, w.theta, nthreads=nthreads)
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        ref.eval(w.u, True, "projected", nthreads=nthreads)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {"value": w.tets.shape[0] / med, "unit": "tets/s", "cores": nthreads, "kind": "port",
            "sample": f"oracle E+g+H(Projected)+CSR on the full 64^3 workload (1,572,864 tets), median of {repeats} "
                      f"evals, {nthreads} threads ({med:.2f} s per eval)"}


def time_steps(dev, step, kind, steps, warmup):
    import torch

    for _ in range(warmup):
        step(kind)
    dev.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(kind)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms, evals = dev.kernel_times()
    return elapsed / steps * 1e3, [m / evals for m in ms]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--exact-only", action="store_true", help="time only the Exact-Hessian step (profiling aid)")
    ap.add_argument("--probe-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a known-size kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the sharded design is described "
              "in DESIGN.md.", file=sys.stderr)
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
        "exact": exact,
    }
    if not args.no_cpu:
        nthreads = min(16, len(os.sched_getaffinity(0)))
        result["cpu_baseline"] = cpu_baseline(w, nthreads, args.cpu_repeats)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
