"""Inputs of the golden fixtures: the key/query sets of the reference's own
tests (tests/test_natinterp3d.py, cited per case) and the configuration-1
workload.  Shared by make_golden.py and the tests that read the fixtures."""
import numpy as np


def _interior(keys, n, rng, margin=0.8):
    # random_interior_queries, tests/test_natinterp3d.py:10-13
    center = keys.mean(axis=0)
    return center + (rng.randn(n, 3)) * keys.std(axis=0) * margin * 0.3


def _cases():
    c = {}
    rng = np.random.RandomState(42)  # :21-30
    k = rng.randn(200, 3)
    c["partition_random"] = (k, _interior(k, 50, rng))
    xs = np.linspace(-1, 1, 6)  # :32-42
    k = np.array(np.meshgrid(xs, xs, xs)).reshape(3, -1).T
    c["partition_grid"] = (k, np.random.RandomState(7).uniform(-0.8, 0.8, (30, 3)))
    rng = np.random.RandomState(99)  # :48-53
    k = rng.randn(300, 3)
    c["nonneg"] = (k, _interior(k, 80, rng))
    rng = np.random.RandomState(12)  # :66-74
    k = rng.randn(500, 3)
    c["linear"] = (k, _interior(k, 60, rng))
    c["cube"] = (np.array([[-1, -1, -1], [-1, -1, 1], [-1, 1, -1], [-1, 1, 1],
                           [1, -1, -1], [1, -1, 1], [1, 1, -1], [1, 1, 1]], float),
                 np.array([[0.0, 0.0, 0.0]]))  # :148-161
    k = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)  # :163-171
    c["regular_tet"] = (k, k.mean(axis=0, keepdims=True))
    c["octahedron"] = (np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1],
                                 [0, 0, -1]], float), np.array([[0.0, 0.0, 0.0]]))  # :173-181
    k = np.random.RandomState(33).randn(50, 3)  # :187-199
    c["data_points"] = (k, k[[0, 10, 25, 49]])
    k = np.random.RandomState(20).randn(100, 3)  # :216-222
    c["far"] = (k, np.array([[100.0, 100.0, 100.0], [-100.0, -100.0, -100.0]]))
    c["five"] = (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0.5, 0.5, 0.5]],
                          float), np.array([[0.2, 0.2, 0.2]]))  # :242-252
    c["four"] = (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float),
                 np.array([[0.1, 0.1, 0.1]]))  # :254-264
    rng = np.random.RandomState(88)  # :282-291
    k = rng.randn(200, 3) + np.array([1e6, -2e6, 3e6])
    c["large_offset"] = (k, _interior(k, 30, rng))
    rng = np.random.RandomState(89)  # :293-301
    k = rng.randn(200, 3) * 1e4
    c["large_scale"] = (k, _interior(k, 30, rng))
    rng = np.random.RandomState(90)  # :303-311
    k = rng.randn(200, 3) * 1e-4
    c["small_scale"] = (k, _interior(k, 30, rng))
    rng = np.random.RandomState(2)  # :348-354
    c["parallel_serial"] = (rng.randn(200, 3), rng.randn(80, 3) * 0.3)
    rng = np.random.RandomState(80)  # :474-484
    c["sparsity"] = (rng.randn(1000, 3), rng.randn(50, 3) * 0.3)
    rng = np.random.RandomState(33)  # edge rows: NaN, inf, duplicates, extrapolation
    k = rng.randn(50, 3)
    c["edge_rows"] = (k, np.vstack([k[[3, 7]], [[3.0, 3, 3], [-2.5, 0, 1], [np.nan, 0, 0],
                                                [np.inf, 0, 0], [0, 0, 0], [0, 0, 0]]]))
    k = np.random.RandomState(5).randn(100, 3)  # duplicate keys
    c["duplicate_keys"] = (np.vstack([k, k[:10]]), np.random.RandomState(6).randn(40, 3) * 0.3)
    return c


CASES = _cases()


def config1():
    """10,000 keys uniform in [0,1]^3 and 100,000 queries in [0.05,0.95]^3
    from numpy default_rng(0)."""
    rng = np.random.default_rng(0)
    keys = rng.random((10_000, 3))
    queries = 0.05 + 0.9 * rng.random((100_000, 3))
    return keys, queries
