"""VEG mesh I/O (litepgo_amd/veg.py) on the CPU: the reference's test assets (box-tet.veg, box-hex.veg, kept as
fixtures under tests/golden/) against an independent line parse, bitwise round trips, *INCLUDE splicing and
cycles, 0-based ids, separators, ignored records, and every error the reader reports
(mirroring tests/python/test_mesh.py of the reference)."""
from pathlib import Path

import numpy as np
import pytest

from litepgo_amd import VegMesh, VolumetricFemEnergy, lame_parameters, read_veg, write_veg

GOLD = Path(__file__).resolve().parent / "golden"


def independent_parse(path):
    lines = [ln.strip() for ln in path.read_text().splitlines() if ln.strip() and not ln.strip().startswith("#")]
    start = lines.index("*VERTICES")
    count = int(lines[start + 1].split()[0])
    vertices = [[float(x) for x in ln.split()[1:4]] for ln in lines[start + 2:start + 2 + count]]
    start = lines.index("*ELEMENTS")
    count = int(lines[start + 2].split()[0])
    cells = [[int(x) - 1 for x in ln.split()[1:]] for ln in lines[start + 3:start + 3 + count]]
    return np.array(vertices), np.array(cells)


@pytest.mark.parametrize("name,width", [("box-tet.veg", 4), ("box-hex.veg", 8)])
def test_reference_assets_against_independent_parse(name, width):
    mesh = read_veg(GOLD / name)
    X, C = independent_parse(GOLD / name)
    assert mesh.positions.dtype == np.float64 and mesh.cells.dtype == np.int32
    assert mesh.cells.shape[1] == width and mesh.element_type == ("TET" if width == 4 else "CUBIC")
    np.testing.assert_array_equal(mesh.positions, X)
    np.testing.assert_array_equal(mesh.cells, C)


def test_box_tet_counts_and_energy():
    mesh = read_veg(GOLD / "box-tet.veg")
    assert (mesh.num_vertices, len(mesh.cells)) == (243, 827)  # the file's own header comment
    lam, mu = lame_parameters(1e5, 0.4)
    energy = VolumetricFemEnergy(mesh.num_vertices, mesh.cells, mesh.positions, np.tile([lam, mu], (827, 1)))
    assert energy.num_dofs == 3 * 243


@pytest.mark.parametrize("name", ["box-tet.veg", "box-hex.veg"])
def test_roundtrip_is_bitwise(tmp_path, name):
    mesh = read_veg(GOLD / name)
    write_veg(tmp_path / name, mesh)
    copy = read_veg(tmp_path / name)
    np.testing.assert_array_equal(copy.positions, mesh.positions)
    np.testing.assert_array_equal(copy.cells, mesh.cells)


def test_precision(tmp_path):
    positions = np.array([[0.12345678901234567, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    write_veg(tmp_path / "m.veg", VegMesh(positions, np.array([[0, 1, 2, 3]], dtype=np.int32)))
    copy = read_veg(tmp_path / "m.veg")
    np.testing.assert_array_equal(copy.positions, positions)
    np.testing.assert_array_equal(copy.cells, [[0, 1, 2, 3]])


def test_include_comments_separators_zero_based_and_ignored_records(tmp_path):
    (tmp_path / "verts.inc").write_text("*VERTICES\n4 3 0 0\n0,0,0,0\n1\t1 0 0\n2 0, 1, 0\n3 0 0 1\n")
    (tmp_path / "m.veg").write_text(
        "# a comment\n\n*INCLUDE verts.inc\n*MATERIAL STVK\nENU, 1000, 1e5, 0.45\n*ELEMENTS\n T E T \n1 4 0\n"
        "7 0 1 2 3\n*SET fixed\n1, 2\n")
    mesh = read_veg(tmp_path / "m.veg")
    np.testing.assert_array_equal(mesh.positions, [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    np.testing.assert_array_equal(mesh.cells, [[0, 1, 2, 3]])


def test_include_cycle(tmp_path):
    (tmp_path / "cycle.veg").write_text("*INCLUDE cycle.veg\n")
    with pytest.raises(RuntimeError, match="cycle"):
        read_veg(tmp_path / "cycle.veg")
    # A longer cycle through a second file, and the same file included twice in sequence (not a cycle).
    (tmp_path / "a.veg").write_text("*INCLUDE b.veg\n")
    (tmp_path / "b.veg").write_text("# b\n*INCLUDE a.veg\n")
    with pytest.raises(RuntimeError, match="cycle: a.veg -> b.veg -> a.veg"):
        read_veg(tmp_path / "a.veg")
    (tmp_path / "v.inc").write_text(GOOD_VERTICES)
    (tmp_path / "e.inc").write_text("*SET s\n1\n")
    (tmp_path / "twice.veg").write_text("*INCLUDE v.inc\n*INCLUDE e.inc\n*INCLUDE e.inc\n*ELEMENTS\nTET\n1 4 0\n"
                                        "1 1 2 3 4\n")
    np.testing.assert_array_equal(read_veg(tmp_path / "twice.veg").cells, [[0, 1, 2, 3]])


GOOD_VERTICES = "*VERTICES\n4 3 0 0\n1 0 0 0\n2 1 0 0\n3 0 1 0\n4 0 0 1\n"


@pytest.mark.parametrize("text,message", [
    ("*VERTICES\n4 2 0 0\n", "3 coordinates per vertex"),
    ("*VERTICES\n5 3 0 0\n1 0 0 0\n", "ends after 1 of 5 rows"),
    ("*VERTICES\n2 3 0 0\n1 0 0 0\n3 1 0 0\n*ELEMENTS\nTET\n0 4 0\n", "consecutive"),
    ("*VERTICES\n1 3 0 0\n7 0 0 0\n*ELEMENTS\nTET\n0 4 0\n", "start at 0 or 1"),
    ("*VERTICES\n1 3 0 0\n1 0 0\n*ELEMENTS\nTET\n0 4 0\n", "'id x y z'"),
    ("*VERTICES\n1 3 0 0\n1 0 x 0\n", "non-numeric vertex row"),
    (GOOD_VERTICES + GOOD_VERTICES, "second \\*VERTICES"),
    (GOOD_VERTICES + "*ELEMENTS\nPRISM\n1 6 0\n", "'PRISM' is not supported"),
    (GOOD_VERTICES + "*ELEMENTS\nTET\n1 8 0\n1 1 2 3 4\n", "TET elements have 4 vertices"),
    (GOOD_VERTICES + "*ELEMENTS\nTET\n2 4 0\n1 1 2 3 4\n", "ends after 1 of 2 rows"),
    (GOOD_VERTICES + "*ELEMENTS\nTET\n1 4 0\n1 1 2 3 5\n", "outside 0..3"),
    (GOOD_VERTICES + "*ELEMENTS\nTET\n1 4 0\n1 1 2 3\n", "4 vertex ids"),
    (GOOD_VERTICES + "*ELEMENTS\n", "element type missing"),
    (GOOD_VERTICES, "no \\*ELEMENTS"),
    ("*ELEMENTS\nTET\n0 4 0\n", "must follow"),
    ("1 2 3\n*VERTICES\n", "outside any section"),
    ("*INCLUDE\n", "without a file name"),
])
def test_errors(tmp_path, text, message):
    (tmp_path / "bad.veg").write_text(text)
    with pytest.raises(RuntimeError, match=message):
        read_veg(tmp_path / "bad.veg")


def test_missing_file(tmp_path):
    with pytest.raises(RuntimeError, match="cannot open"):
        read_veg(tmp_path / "absent.veg")
