"""Host-side logic of the multi-mesh batch (no GPU): offsets and the block-diagonal union pattern."""
import numpy as np
import pytest

from litepgo_amd import BatchedFemEnergy, VolumetricFemEnergy, workloads


def test_union_pattern_is_block_diagonal_of_the_meshes():
    meshes = workloads.batched(num_meshes=5, seed=3, lo=2, hi=4)
    batch = BatchedFemEnergy(meshes)
    vo, to, no = batch.offsets()
    assert batch.num_meshes == 5 and vo[-1] == batch.num_vertices and to[-1] == batch.num_tets
    row_ptr, col_idx = batch.pattern()
    assert no[-1] == batch.nnz == row_ptr[-1]
    for m, w in enumerate(meshes):
        plain = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
        r, c = plain.pattern()
        d0, d1 = 3 * vo[m], 3 * vo[m + 1]
        assert np.array_equal(row_ptr[d0:d1 + 1] - row_ptr[d0], r)
        assert np.array_equal(col_idx[no[m]:no[m + 1]], c + d0)
        assert row_ptr[d0] == no[m]


def test_batch_construction_errors():
    w = workloads.batched(num_meshes=2, seed=1, lo=2, hi=2)
    bad = workloads.batched(num_meshes=1, seed=2, lo=2, hi=2)[0]
    tets = bad.tets.copy()
    tets[0, 0] = bad.num_vertices  # outside the mesh's own vertex range
    with pytest.raises(ValueError, match="outside"):
        BatchedFemEnergy([w[0], (bad.num_vertices, tets, bad.rest, bad.theta)])
    with pytest.raises(ValueError):
        BatchedFemEnergy([])
