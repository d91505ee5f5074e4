"""CPU test of the 9x9 symmetric eigen-solver that the PSD-projection kernel runs per indefinite element
(litepgo_amd/csrc/lpg_eig9.hpp, compiled here as plain C++): orthonormal eigenvectors, reconstruction of the
matrix, and eigenvalues against an independent cyclic Jacobi, on random, degenerate and neo-Hookean matrices; and the
PSD projection on matrices with exactly repeated eigenvalues kept as a fixture (tests/golden/eig9_repeated_eigenvalues.txt)."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_eig9_solver(tmp_path):
    exe = tmp_path / "eig9_test"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-Wno-unknown-pragmas", "-ffp-contract=off",
                    f"-I{ROOT / 'litepgo_amd' / 'csrc'}", str(ROOT / "tests" / "cpp" / "eig9_test.cpp"),
                    "-o", str(exe)], check=True)
    fixture = ROOT / "tests" / "golden" / "eig9_repeated_eigenvalues.txt"
    out = subprocess.run([str(exe), str(fixture)], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
