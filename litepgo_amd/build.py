"""Builds the HIP library litepgo_amd/liblitepgo_fem.so for gfx950 with hipcc (in-tree, no JIT cache)."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "liblitepgo_fem.so"
SOURCES = ["lpg_kernels.hip", "lpg_solver.hip", "lpg_setup.cpp", "lpg_api.cpp", "lpg_batch.cpp", "lpg_newton.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops stays off (hand-written FP64 code only; packed FP32 forms are avoided on these parts).
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-munsafe-fp-atomics",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          f"-I{ROOT / 'include'}", f"-I{CSRC}", "-Wall", "-Wno-unused-result"]


def build(verbose: bool = False, extra: list[str] | None = None) -> Path:
    """Compile every source and link the shared library; always rebuilds."""
    from concurrent.futures import ThreadPoolExecutor

    build_dir = PKG / "_build"
    build_dir.mkdir(exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        obj = build_dir / (src + ".o")
        cmd = [HIPCC, *COMMON, *(extra or []), "-c", str(CSRC / src), "-o", str(obj)]
        if src.endswith(".cpp"):
            cmd.insert(1, "-x")
            cmd.insert(2, "hip")
        if verbose:
            print(" ".join(cmd), flush=True)
        cmds.append(cmd)
        objs.append(str(obj))
    # Sources compile independently; at most four at a time.
    with ThreadPoolExecutor(max_workers=min(4, len(cmds))) as pool:
        for proc in pool.map(lambda c: subprocess.run(c, check=True), cmds):
            pass
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
