"""The C-ABI library: it loads, exports every function include/*.h declares, and the ctypes table matches."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

from litepgo_amd import _native

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        for m in re.finditer(r"\b(lpg_[A-Za-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_headers_declare_the_ctypes_table():
    assert declared_functions() == set(_native.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_native.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert declared_functions() <= exported


def test_headers_compile_as_c_and_cpp(tmp_path):
    src = tmp_path / "use.c"
    src.write_text('#include "litepgo_fem.h"\n#include "litepgo_fem_ext.h"\n#include "litepgo_fem_batch.h"\nint main(void) { return LPG_OK; }\n')
    for compiler, name in (("gcc", "use.c"), ("g++", "use.cpp")):
        target = tmp_path / name
        target.write_text(src.read_text())
        subprocess.run([compiler, "-std=c11" if compiler == "gcc" else "-std=c++17", "-Wall", "-Werror",
                        f"-I{ROOT / 'include'}", "-c", str(target), "-o", str(tmp_path / (name + ".o"))],
                       check=True)
