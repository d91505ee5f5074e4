"""VEG volume-mesh files: read and write, for feeding real tetrahedral meshes to the device operator.

Follows litepgo's reader and writer (src/core/mesh/io/veg.cpp:1-235; Python binding src/binding/mesh/io/veg.cpp):

* lines are trimmed; blank lines and lines starting with '#' are skipped; ``*INCLUDE <file>`` splices in another file
  (relative to the including file), and an include cycle is an error;
* ``*VERTICES``: a header ``count 3 ...``, then ``count`` lines ``id x y z`` whose ids run consecutively from 0 or 1
  (the first line decides);
* ``*ELEMENTS``: after exactly one vertex section, a type line ``TET`` (4 vertices) or ``CUBIC`` (8), a header
  ``count width ...`` whose width matches the type, then ``count`` lines ``id v_1 .. v_width`` with vertex ids in range;
* other records (materials, sets, regions) are ignored; commas and tabs separate numbers like spaces;
* malformed or truncated data raise ``RuntimeError``, as the reference's ``std::runtime_error`` does through its
  binding.

The writer emits 1-based ids and 17 significant digits, so a round trip reproduces positions bit for bit.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

__all__ = ["VegMesh", "read_veg", "write_veg"]

_WIDTH = {"TET": 4, "CUBIC": 8}


@dataclass
class VegMesh:
    """positions (V, 3) float64; cells (C, 4) for TET or (C, 8) for CUBIC, int32, 0-based."""

    positions: np.ndarray
    cells: np.ndarray

    @property
    def element_type(self) -> str:
        return "TET" if self.cells.shape[1] == 4 else "CUBIC"

    @property
    def num_vertices(self) -> int:
        return int(self.positions.shape[0])


def _lines(path: Path, chain: tuple = ()) -> list[str]:
    canonical = Path(os.path.realpath(path))
    if canonical in chain:
        raise RuntimeError(f"veg *INCLUDE cycle through {path}")
    chain = chain + (canonical,)
    try:
        text = Path(path).read_text()
    except OSError as exc:
        raise RuntimeError(f"Failed to open veg file: {path}") from exc
    out = []
    for raw in text.splitlines():
        line = raw.strip()
        if not line or line.startswith("#"):
            continue
        if line.startswith("*INCLUDE "):
            out.extend(_lines(Path(path).parent / line[len("*INCLUDE "):].strip(), chain))
            continue
        out.append(line)
    return out


def _numbers(line: str, kind, path: Path) -> list:
    fields = line.replace(",", " ").replace("\t", " ").split()
    try:
        return [kind(f) for f in fields]
    except ValueError as exc:
        raise RuntimeError(f"Malformed VEG numeric record in {path}: {line}") from exc


def read_veg(path) -> VegMesh:
    path = Path(path)
    lines = _lines(path)
    vertices: list | None = None
    cells: list | None = None
    width = 0
    base = 1
    i = 0
    while i < len(lines):
        line = lines[i]
        if line.startswith("*VERTICES"):
            if vertices is not None:
                raise RuntimeError(f"Duplicate *VERTICES section in {path}")
            i += 1
            if i >= len(lines):
                raise RuntimeError(f"Missing *VERTICES header in {path}")
            header = _numbers(lines[i], int, path)
            i += 1
            if len(header) < 2 or header[0] < 0 or header[1] != 3:
                raise RuntimeError(f"Invalid 3D vertex header in {path}")
            count = header[0]
            if count > len(lines) - i:
                raise RuntimeError(f"Truncated vertex data in {path}")
            vertices = []
            for row in range(count):
                values = _numbers(lines[i], float, path)
                if len(values) < 4:
                    raise RuntimeError(f"Malformed vertex line in {path}: {lines[i]}")
                if row == 0 and values[0] == 0:
                    base = 0
                if values[0] != row + base:
                    raise RuntimeError(f"Vertex IDs must be consecutive in {path}")
                vertices.append(values[1:4])
                i += 1
            continue
        if line.startswith("*ELEMENTS"):
            if vertices is None or width:
                raise RuntimeError(f"Expected one *VERTICES section followed by one *ELEMENTS section in {path}")
            i += 1
            if i >= len(lines):
                raise RuntimeError(f"Missing element type in {path}")
            kind = "".join(lines[i].split())
            i += 1
            if kind not in _WIDTH:
                raise RuntimeError(f"Unsupported element type in {path}: {kind}")
            width = _WIDTH[kind]
            if i >= len(lines):
                raise RuntimeError(f"Missing element count in {path}")
            header = _numbers(lines[i], int, path)
            i += 1
            if not header or header[0] < 0:
                raise RuntimeError(f"Invalid element count in {path}")
            count = header[0]
            if count > len(lines) - i:
                raise RuntimeError(f"Truncated element data in {path}")
            if len(header) < 2 or header[1] != width:
                raise RuntimeError(f"Element width does not match its type in {path}")
            cells = []
            for _ in range(count):
                values = _numbers(lines[i], int, path)
                if len(values) < width + 1:
                    raise RuntimeError(f"Malformed element line in {path}: {lines[i]}")
                ids = values[1:width + 1]
                if any(v < base or v - base >= len(vertices) for v in ids):
                    raise RuntimeError(f"Element vertex index out of range in {path}")
                cells.append([v - base for v in ids])
                i += 1
            continue
        i += 1  # legacy material, set, region and other records
    if not width:
        raise RuntimeError(f"No *ELEMENTS section found in {path}")
    positions = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
    if not np.all(np.isfinite(positions)):
        raise RuntimeError(f"Non-finite vertex position in {path}")
    return VegMesh(positions, np.asarray(cells, dtype=np.int32).reshape(-1, width))


def write_veg(path, mesh: VegMesh) -> None:
    positions = np.asarray(mesh.positions, dtype=np.float64).reshape(-1, 3)
    cells = np.asarray(mesh.cells).reshape(len(mesh.cells), -1)
    width = cells.shape[1]
    if width not in (4, 8):
        raise ValueError("write_veg: cells must have 4 (TET) or 8 (CUBIC) vertices")
    if cells.size and (cells.min() < 0 or cells.max() >= len(positions)):
        raise ValueError("write_veg: cell vertex index out of range")
    if not np.all(np.isfinite(positions)):
        raise ValueError("write_veg: non-finite vertex position")
    out = [f"*VERTICES\n{len(positions)} 3 0 0\n"]
    out.extend(f"{v + 1} {p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n" for v, p in enumerate(positions))
    out.append(f"*ELEMENTS\n{'TET' if width == 4 else 'CUBIC'}\n{len(cells)} {width} 0\n")
    out.extend(f"{e + 1} " + " ".join(str(int(v) + 1) for v in c) + "\n" for e, c in enumerate(cells))
    try:
        Path(path).write_text("".join(out))
    except OSError as exc:
        raise RuntimeError(f"Cannot write VEG: {path}") from exc
