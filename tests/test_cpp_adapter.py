"""The header-only C++ adapter (include/litepgo_fem.hpp): it compiles and links against the library with g++
(CPU), and its test program passes on the GPU."""
import subprocess
from pathlib import Path

import pytest

from litepgo_amd import _native

ROOT = Path(__file__).resolve().parents[1]


def build_test(out_dir):
    exe = Path(out_dir) / "energy_adapter_test"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "cpp" / "energy_adapter_test.cpp"), "-o", str(exe),
                    f"-L{_native.LIB_PATH.parent}", "-llitepgo_fem", f"-Wl,-rpath,{_native.LIB_PATH.parent}"],
                   check=True)
    return exe


def test_adapter_compiles_and_links(tmp_path):
    assert build_test(tmp_path).exists()


@pytest.mark.gpu
def test_adapter_program_passes_on_gpu(tmp_path):
    exe = build_test(tmp_path)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stdout + out.stderr
