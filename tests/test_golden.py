"""The oracle against the committed golden fixtures (tests/golden/make_golden.py: an independent NumPy/LAPACK/SciPy
restatement).  Tolerances: 1e-12 relative max-norm per output array; integers bit-exact."""
import hashlib
from pathlib import Path

import numpy as np

from litepgo_amd.fem import kuhn_grid, lame_parameters

GOLD = Path(__file__).resolve().parent / "golden"


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_two_tet_fixture(oracle_mod):
    d = np.load(GOLD / "two_tet.npz")
    E = oracle_mod.OracleEnergy(5, d["tets"], d["X"], d["theta"])
    o, i = E.pattern()
    assert np.array_equal(o, d["outer"]) and np.array_equal(i, d["inner"])
    assert np.array_equal(E.offsets(), d["offsets"])
    for s in range(3):
        v, g, H = E.eval(d["x"][s], True, "exact")
        _, _, Hp = E.eval(d["x"][s], False, "projected")
        assert abs(v - d["value_exact"][s]) <= 1e-12 * max(abs(d["value_exact"][s]), 1e-300)
        if s:
            assert rel(g, d["grad"][s]) <= 1e-12
        else:
            assert np.abs(g).max() == 0.0
        assert rel(H, d["H_exact"][s]) <= 1e-12
        assert rel(Hp, d["H_projected"][s]) <= 1e-12


def test_grid2_fixture(oracle_mod):
    d = np.load(GOLD / "grid2.npz")
    E = oracle_mod.OracleEnergy(len(d["X"]), d["tets"], d["X"], d["theta"])
    o, i = E.pattern()
    assert E.nnz == 2007
    assert np.array_equal(o, d["outer"]) and np.array_equal(i, d["inner"])
    assert np.array_equal(E.offsets(), d["offsets"])
    for s in range(2):
        v, g, H = E.eval(d["x"][s], True, "exact")
        _, _, Hp = E.eval(d["x"][s], False, "projected")
        assert abs(v - d["value_exact"][s]) <= 1e-12 * abs(d["value_exact"][s])
        assert rel(g, d["grad"][s]) <= 1e-12
        assert rel(H, d["H_exact"][s]) <= 1e-12
        assert rel(Hp, d["H_projected"][s]) <= 1e-12
        # the compressed state exercises the projection: it must change H
        if s == 1:
            assert rel(Hp, H) > 1e-6


def test_grid10_fixture(oracle_mod):
    d = np.load(GOLD / "grid10.npz")
    X, tets = kuhn_grid(10)
    lam, mu = lame_parameters(1e5, 0.4)
    E = oracle_mod.OracleEnergy(len(X), tets, X, np.tile([[float(lam), float(mu)]], (len(tets), 1)))
    o, i = E.pattern()
    assert E.nnz == int(d["nnz"]) == 154719
    assert sha(o) == str(d["outer_sha"]) and sha(i) == str(d["inner_sha"])
    assert sha(E.offsets()) == str(d["offsets_sha"])
    v, g, H = E.eval(d["x"], True, "exact")
    _, _, Hp = E.eval(d["x"], False, "projected")
    s = d["sample"]
    assert abs(v - d["value_exact"]) <= 1e-12 * abs(d["value_exact"])
    assert rel(g, d["grad"]) <= 1e-12
    assert np.abs(H[s] - d["H_exact_sample"]).max() <= 1e-12 * d["H_exact_absmax"]
    assert np.abs(Hp[s] - d["H_projected_sample"]).max() <= 1e-12 * d["H_projected_absmax"]
