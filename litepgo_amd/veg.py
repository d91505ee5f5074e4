"""VEG volume-mesh files (the Vega FEM text format litepgo reads, src/core/mesh/io/veg.cpp): read and write, so that
real tetrahedral meshes can be fed to the device operator.

The format, as this reader accepts it:

* a file is a sequence of records, one per line; blank lines and ``#`` comments carry nothing, and commas or tabs
  separate numbers like spaces;
* a line starting with ``*`` opens a section named by its first word; ``*INCLUDE <file>`` is not a section but is
  replaced by the records of that file (a path relative to the including file); including a file that is already
  being read is an error (a cycle);
* ``*VERTICES``: a header ``count 3 ...`` and ``count`` rows ``id x y z``; ids run consecutively from the first
  row's id, which must be 0 or 1 and sets the base of every index in the file;
* ``*ELEMENTS``: the element type (``TET``, 4 vertices, or ``CUBIC``, 8; spaces inside the word are ignored), a
  header ``count width ...`` with the width the type implies, and ``count`` rows ``id v_1 .. v_width``;
* every other section (materials, sets, regions, ...) is skipped;
* a file holds exactly one vertex section, and an element section after it.

Anything else raises ``RuntimeError`` (the reference's std::runtime_error as its binding maps it; its tests pin the
class and the word "cycle": tests/python/test_mesh.py:46-50,146-150).  The writer emits 1-based ids and 17
significant digits, so writing and reading back reproduces positions bit for bit.

Design: records come from a generator that expands includes depth-first with the chain of open files; a small
state machine groups them into sections; each known section is then decoded from its own record list.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass
from pathlib import Path

import numpy as np

__all__ = ["VegMesh", "read_veg", "write_veg"]

_ELEMENT_WIDTH = {"TET": 4, "CUBIC": 8}
_SPLIT = re.compile(r"[,\s]+")


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


@dataclass
class _Record:
    text: str
    origin: str  # "file:line", for messages


def _records(path: Path, open_chain: tuple):
    """Yields the file's records in order with *INCLUDE expanded; open_chain holds the files being read."""
    key = os.path.realpath(path)
    if key in open_chain:
        names = " -> ".join(os.path.basename(p) for p in open_chain + (key,))
        raise RuntimeError(f"VEG include cycle: {names}")
    try:
        text = Path(path).read_text()
    except OSError as exc:
        raise RuntimeError(f"VEG: cannot open {path}") from exc
    chain = open_chain + (key,)
    for number, raw in enumerate(text.splitlines(), start=1):
        line = raw.strip()
        if not line or line.startswith("#"):
            continue
        head = line.split(None, 1)
        if head[0].upper() == "*INCLUDE":
            if len(head) < 2:
                raise RuntimeError(f"VEG: *INCLUDE without a file name at {path}:{number}")
            yield from _records(Path(path).parent / head[1].strip(), chain)
        else:
            yield _Record(line, f"{path}:{number}")


def _sections(records):
    """Groups records into (NAME, header record, body records) in file order; records before any section are an
    error."""
    current = None
    for rec in records:
        if rec.text.startswith("*"):
            if current is not None:
                yield current
            current = (rec.text[1:].split(None, 1)[0].upper() if len(rec.text) > 1 else "", rec, [])
        elif current is None:
            raise RuntimeError(f"VEG: data outside any section at {rec.origin}")
        else:
            current[2].append(rec)
    if current is not None:
        yield current


def _fields(rec: _Record, parse, what: str) -> list:
    try:
        return [parse(tok) for tok in _SPLIT.split(rec.text) if tok]
    except ValueError as exc:
        raise RuntimeError(f"VEG: non-numeric {what} at {rec.origin}: {rec.text!r}") from exc


def _counted_rows(body, header_index: int, header_what: str, check_header):
    """(header numbers, the rows the header's count announces) of a section body; check_header(header, record)
    validates the header before its rows are taken."""
    if len(body) <= header_index:
        raise RuntimeError(f"VEG: {header_what} header missing")
    header = _fields(body[header_index], int, f"{header_what} header")
    if not header or header[0] < 0:
        raise RuntimeError(f"VEG: bad {header_what} header at {body[header_index].origin}")
    check_header(header, body[header_index])
    rows = body[header_index + 1:header_index + 1 + header[0]]
    if len(rows) < header[0]:
        raise RuntimeError(f"VEG: {header_what} section ends after {len(rows)} of {header[0]} rows")
    return header, rows


def _decode_vertices(body):
    def check(header, rec):
        if len(header) < 2 or header[1] != 3:
            raise RuntimeError(f"VEG: vertex header must give 3 coordinates per vertex ({rec.origin})")

    header, rows = _counted_rows(body, 0, "vertex", check)
    positions = np.empty((len(rows), 3), dtype=np.float64)
    base = None
    for k, rec in enumerate(rows):
        vals = _fields(rec, float, "vertex row")
        if len(vals) != 4 or not vals[0].is_integer():
            raise RuntimeError(f"VEG: a vertex row is 'id x y z' ({rec.origin}: {rec.text!r})")
        if base is None:
            base = int(vals[0])
            if base not in (0, 1):
                raise RuntimeError(f"VEG: vertex ids must start at 0 or 1 ({rec.origin})")
        if int(vals[0]) != base + k:
            raise RuntimeError(f"VEG: vertex ids must be consecutive ({rec.origin})")
        positions[k] = vals[1:]
    if not np.isfinite(positions).all():
        raise RuntimeError("VEG: non-finite vertex coordinate")
    return positions, (1 if base is None else base)


def _decode_elements(body, num_vertices: int, base: int):
    if not body:
        raise RuntimeError("VEG: element type missing")
    kind = "".join(body[0].text.split()).upper()
    if kind not in _ELEMENT_WIDTH:
        raise RuntimeError(f"VEG: element type {kind!r} is not supported (TET or CUBIC) ({body[0].origin})")
    width = _ELEMENT_WIDTH[kind]

    def check(header, rec):
        if len(header) < 2 or header[1] != width:
            raise RuntimeError(f"VEG: {kind} elements have {width} vertices, the header says "
                               f"{header[1] if len(header) > 1 else '?'} ({rec.origin})")

    header, rows = _counted_rows(body, 1, "element", check)
    cells = np.empty((len(rows), width), dtype=np.int64)
    for k, rec in enumerate(rows):
        vals = _fields(rec, int, "element row")
        if len(vals) != width + 1:
            raise RuntimeError(f"VEG: an element row is 'id' and {width} vertex ids ({rec.origin}: {rec.text!r})")
        cells[k] = vals[1:]
    cells -= base
    if cells.size and (cells.min() < 0 or cells.max() >= num_vertices):
        raise RuntimeError(f"VEG: element vertex id outside 0..{num_vertices - 1} (after the file's base {base})")
    return cells.astype(np.int32)


def read_veg(path) -> VegMesh:
    """Reads a VEG file (tet or hex elements) into a VegMesh with 0-based cells."""
    vertices = elements = None
    for name, rec, body in _sections(_records(Path(path), ())):
        if name == "VERTICES":
            if vertices is not None:
                raise RuntimeError(f"VEG: a second *VERTICES section at {rec.origin}")
            vertices = _decode_vertices(body)
        elif name == "ELEMENTS":
            if vertices is None or elements is not None:
                raise RuntimeError(f"VEG: *ELEMENTS must follow the one *VERTICES section ({rec.origin})")
            elements = _decode_elements(body, vertices[0].shape[0], vertices[1])
    if vertices is None or elements is None:
        raise RuntimeError(f"VEG: {path} has no {'*VERTICES' if vertices is None else '*ELEMENTS'} section")
    return VegMesh(vertices[0], elements)


def write_veg(path, mesh: VegMesh) -> None:
    """Writes positions and cells as a VEG file (1-based ids, 17 significant digits)."""
    positions = np.asarray(mesh.positions, dtype=np.float64).reshape(-1, 3)
    cells = np.asarray(mesh.cells, dtype=np.int64)
    if cells.ndim != 2 or cells.shape[1] not in (4, 8):
        raise ValueError("write_veg: cells must have 4 (TET) or 8 (CUBIC) vertices")
    if cells.size and (cells.min() < 0 or cells.max() >= positions.shape[0]):
        raise ValueError("write_veg: cell vertex index out of range")
    if not np.isfinite(positions).all():
        raise ValueError("write_veg: non-finite vertex position")
    kind = "TET" if cells.shape[1] == 4 else "CUBIC"
    out = [f"# {positions.shape[0]} vertices, {cells.shape[0]} elements", "", "*VERTICES",
           f"{positions.shape[0]} 3 0 0"]
    out += [f"{i} {x:.17g} {y:.17g} {z:.17g}" for i, (x, y, z) in enumerate(positions.tolist(), start=1)]
    out += ["", "*ELEMENTS", kind, f"{cells.shape[0]} {cells.shape[1]} 0"]
    out += [f"{i} " + " ".join(str(v + 1) for v in row) for i, row in enumerate(cells.tolist(), start=1)]
    try:
        Path(path).write_text("\n".join(out) + "\n")
    except OSError as exc:
        raise RuntimeError(f"VEG: cannot write {path}") from exc
