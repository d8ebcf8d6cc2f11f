"""Mesh reuse: a mesh saved to disk and loaded back (no builder run), and the
module-level calls' mesh cache, give rows bit-identical to a fresh build."""
import numpy as np
import pytest

import natinterp3d_amd
from natinterp3d_amd import workloads

pytestmark = pytest.mark.gpu


def csr_equal(a, b):
    return (np.array_equal(a.indptr, b.indptr) and np.array_equal(a.indices, b.indices)
            and np.array_equal(a.data, b.data))


def test_saved_interpolator_gives_the_same_rows(tmp_path):
    k, q, v = workloads.uniform_small(seed=21, n_keys=5000, n_queries=4000)
    it = natinterp3d_amd.Interpolator(k)
    W = it.get_weights(q)
    y = it.interpolate(q, v)
    path = tmp_path / "interp.npz"
    it.save(path)
    it2 = natinterp3d_amd.Interpolator.load(path)
    assert csr_equal(it2.get_weights(q), W)
    assert np.array_equal(it2.interpolate(q, v), y)
    assert it2.mesh_and_vertices.info()["n_tets"] == it.mesh_and_vertices.info()["n_tets"]
    # a shell (rows from the larger levels) through the mesh-only form
    ks, qs, _, _ = workloads.shell(seed=3, n_keys=1500, n_queries=300)
    its = natinterp3d_amd.Interpolator(ks)
    its.mesh_and_vertices.save(tmp_path / "mesh.npz")
    m2 = natinterp3d_amd.MeshAndVerticesHIP.load(tmp_path / "mesh.npz")
    qc = np.ascontiguousarray(qs - its._centroid)
    assert csr_equal(m2.get_weights(qc), its.mesh_and_vertices.get_weights(qc))


def test_module_level_calls_reuse_the_mesh():
    k, q, v = workloads.uniform_small(seed=22, n_keys=3000, n_queries=2000)
    natinterp3d_amd.clear_mesh_cache()
    s0 = natinterp3d_amd.mesh_cache_stats()
    W1 = natinterp3d_amd.get_weights(q, k)
    y1 = natinterp3d_amd.interpolate(q, k, v)
    W2 = natinterp3d_amd.get_weights(q[::-1], k.copy())  # same keys, new array
    s1 = natinterp3d_amd.mesh_cache_stats()
    assert s1["misses"] - s0["misses"] == 1 and s1["hits"] - s0["hits"] == 2
    fresh = natinterp3d_amd.Interpolator(k)
    assert csr_equal(W1, fresh.get_weights(q))
    assert csr_equal(W2, fresh.get_weights(q[::-1]))
    assert np.array_equal(y1, fresh.interpolate(q, v))
    # other keys: a new mesh; the cache keeps at most two
    k2 = k + 1e-3
    natinterp3d_amd.get_weights(q, k2)
    natinterp3d_amd.get_weights(q, k2 * 2.0)
    s2 = natinterp3d_amd.mesh_cache_stats()
    assert s2["entries"] == 2 and s2["misses"] - s1["misses"] == 2
    natinterp3d_amd.clear_mesh_cache()
    assert natinterp3d_amd.mesh_cache_stats()["entries"] == 0
