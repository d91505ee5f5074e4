"""CPU test of the FP64 logarithm the element kernels use for log J (litepgo_amd/csrc/lpg_math.hpp, built here as
plain C++ with the host division in place of the device reciprocal): within 2.5 ulp of std::log everywhere tested,
NaN outside (0, inf)."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_log_pos_accuracy(tmp_path):
    exe = tmp_path / "math_test"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-ffp-contract=off",
                    f"-I{ROOT / 'litepgo_amd' / 'csrc'}", str(ROOT / "tests" / "cpp" / "math_test.cpp"),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
