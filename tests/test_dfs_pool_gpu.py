"""The DFS kernel keeps a wave's in-sphere adjacency in one LDS pool and
walks a query whose words do not fit there in its record instead (a second
copy of the walk over global memory).  The test hook
NATINTERP3D_TEST_DFS_POOL shrinks the pool, so that every walk (0) or the
later lanes of each wave (700 words) take that copy: the rows must be the
same, bit for bit, as with the whole pool (which the parity tests pin to
the oracle)."""
import numpy as np
import pytest

from natinterp3d_amd import workloads
from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mesh_and_queries():
    keys, queries, _ = workloads.plain_at_scale(n_keys=20000, n_queries=200000)
    c = keys.mean(axis=0)
    m = MeshAndVerticesHIP(np.ascontiguousarray(keys - c), device=0)
    return m, np.ascontiguousarray(queries - c)


@pytest.mark.parametrize("words", ["0", "700"])
def test_rows_without_the_lds_pool_are_identical(mesh_and_queries, monkeypatch, words):
    m, q = mesh_and_queries
    monkeypatch.delenv("NATINTERP3D_TEST_DFS_POOL", raising=False)
    ref = m.get_weights(q)
    monkeypatch.setenv("NATINTERP3D_TEST_DFS_POOL", words)
    got = m.get_weights(q)
    assert np.array_equal(got.indptr, ref.indptr)
    assert np.array_equal(got.indices, ref.indices)
    assert np.array_equal(got.data, ref.data)
