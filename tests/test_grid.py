"""kuhn_grid against a literal restatement of benchmarks/contact_scaling.py:17-36 (bitwise)."""
import itertools
import math

import numpy as np
import pytest

from litepgo_amd.fem import kuhn_grid


def reference_loop(ns, bounds):
    nx, ny, nz = ns
    vertices = list(itertools.product(*([lo + (hi - lo) * i / n for i in range(n + 1)]
                                        for (lo, hi), n in zip(bounds, ns))))

    def index(p):
        return (p[0] * (ny + 1) + p[1]) * (nz + 1) + p[2]

    tets = []
    for cell in itertools.product(range(nx), range(ny), range(nz)):
        for perm in itertools.permutations(range(3)):
            p = list(cell)
            tet = [index(p)]
            for axis in perm:
                p[axis] += 1
                tet.append(index(p))
            a, b, c, d = (vertices[i] for i in tet)
            u, v, w = ([q[i] - a[i] for i in range(3)] for q in (b, c, d))
            det = sum(u[i] * (v[(i + 1) % 3] * w[(i + 2) % 3] - v[(i + 2) % 3] * w[(i + 1) % 3]) for i in range(3))
            if det < 0:
                tet[1], tet[2] = tet[2], tet[1]
                det = -det
            assert det > 0
            tets.append(tet)
    return np.array(vertices), np.array(tets, dtype=np.int32)


@pytest.mark.parametrize("ns,bounds", [((1, 1, 1), [(0, 1)] * 3), ((3, 3, 3), [(0, 1)] * 3),
                                       ((2, 3, 4), [(0.0, 0.2), (0.0, 0.3), (0.0, 0.4)]),
                                       ((5, 5, 5), [(-1.0, 2.0), (0.5, 0.75), (0, 1)])])
def test_matches_reference_generator(ns, bounds):
    Xr, Tr = reference_loop(ns, bounds)
    X, T = kuhn_grid(ns, bounds)
    assert np.array_equal(X, Xr)  # bitwise
    assert np.array_equal(T, Tr)


def test_counts_and_volume():
    for n in (1, 4, 10):
        X, T = kuhn_grid(n)
        assert len(T) == 6 * n ** 3 and len(X) == (n + 1) ** 3
        D = np.stack([X[T[:, k + 1]] - X[T[:, 0]] for k in range(3)], axis=2)
        vol = np.linalg.det(D) / 6
        assert np.all(vol > 0)
        assert math.isclose(vol.sum(), 1.0, rel_tol=1e-11)
