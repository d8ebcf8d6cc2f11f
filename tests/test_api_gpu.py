"""The reference's API test-suite (tests/test_natinterp3d.py, 39 tests in 18
classes), run against natinterp3d_amd: same inputs, same assertions.  Every
weight here comes from the HIP kernels."""
import numpy as np
import pytest
import scipy.sparse

import natinterp3d_amd as natinterp3d

pytestmark = pytest.mark.gpu


def random_interior_queries(keys, n_queries, rng, margin=0.8):
    center = keys.mean(axis=0)
    return center + (rng.randn(n_queries, 3)) * keys.std(axis=0) * margin * 0.3


class TestPartitionOfUnity:  # :18-42
    def test_random_queries(self):
        rng = np.random.RandomState(42)
        keys = rng.randn(200, 3)
        queries = random_interior_queries(keys, 50, rng)
        W = natinterp3d.get_weights(queries, keys)
        sums = np.asarray(W.sum(axis=1)).ravel()
        interior = sums > 0.5
        assert interior.sum() > 30
        np.testing.assert_allclose(sums[interior], 1.0, atol=1e-10)

    def test_grid_data(self):
        xs = np.linspace(-1, 1, 6)
        keys = np.array(np.meshgrid(xs, xs, xs)).reshape(3, -1).T
        rng = np.random.RandomState(7)
        queries = rng.uniform(-0.8, 0.8, (30, 3))
        W = natinterp3d.get_weights(queries, keys)
        sums = np.asarray(W.sum(axis=1)).ravel()
        interior = sums > 0.5
        assert interior.sum() > 20
        np.testing.assert_allclose(sums[interior], 1.0, atol=1e-10)


class TestNonNegativity:  # :45-60
    def test_random(self):
        rng = np.random.RandomState(99)
        keys = rng.randn(300, 3)
        queries = random_interior_queries(keys, 80, rng)
        W = natinterp3d.get_weights(queries, keys)
        assert W.min() >= -1e-12

    def test_serial(self):
        rng = np.random.RandomState(99)
        keys = rng.randn(300, 3)
        queries = random_interior_queries(keys, 80, rng)
        W = natinterp3d.get_weights(queries, keys, parallel=False)
        assert W.min() >= -1e-12


class TestLinearPrecision:  # :63-99
    @pytest.fixture()
    def setup(self):
        rng = np.random.RandomState(12)
        keys = rng.randn(500, 3)
        queries = random_interior_queries(keys, 60, rng)
        W = natinterp3d.get_weights(queries, keys)
        sums = np.asarray(W.sum(axis=1)).ravel()
        return keys, queries, W, sums > 0.5

    def test_constant_function(self, setup):
        keys, queries, W, interior = setup
        result = np.asarray(W @ np.full(len(keys), 7.0)[:, None]).ravel()
        np.testing.assert_allclose(result[interior], 7.0, atol=1e-10)

    def test_linear_xyz(self, setup):
        keys, queries, W, interior = setup
        coeffs = np.array([2.0, -3.0, 5.0])
        values = keys @ coeffs + 1.0
        expected = queries @ coeffs + 1.0
        result = np.asarray(W @ values[:, None]).ravel()
        np.testing.assert_allclose(result[interior], expected[interior], atol=1e-8)

    def test_each_coordinate(self, setup):
        keys, queries, W, interior = setup
        for dim in range(3):
            result = np.asarray(W @ keys[:, dim][:, None]).ravel()
            np.testing.assert_allclose(result[interior], queries[interior, dim], atol=1e-8)


class TestLocalCoordinateProperty:  # :102-114
    def test_weighted_position(self):
        rng = np.random.RandomState(55)
        keys = rng.randn(300, 3)
        queries = random_interior_queries(keys, 40, rng)
        W = natinterp3d.get_weights(queries, keys)
        interior = np.asarray(W.sum(axis=1)).ravel() > 0.5
        errors = np.linalg.norm(np.asarray(W @ keys)[interior] - queries[interior], axis=1)
        np.testing.assert_allclose(errors, 0.0, atol=1e-8)


class TestConvergence:  # :117-140
    def test_smooth_function(self):
        rng = np.random.RandomState(10)

        def func(x):
            return np.sin(x[:, 0]) + np.cos(x[:, 1]) + 0.5 * x[:, 2]

        queries = rng.randn(100, 3) * 0.3
        errors = []
        for n in [200, 2000, 20000]:
            keys = rng.randn(n, 3)
            result = natinterp3d.interpolate(queries, keys, func(keys))
            errors.append(np.mean(np.abs(result - func(queries))))
        assert errors[1] < errors[0]
        assert errors[2] < errors[1]
        assert errors[2] < 2e-3


class TestSymmetricConfigurations:  # :145-181
    def test_cube_center(self):
        keys = np.array([[-1, -1, -1], [-1, -1, 1], [-1, 1, -1], [-1, 1, 1],
                         [1, -1, -1], [1, -1, 1], [1, 1, -1], [1, 1, 1]], dtype=np.float64)
        W = natinterp3d.get_weights(np.array([[0.0, 0.0, 0.0]]), keys)
        np.testing.assert_allclose(np.asarray(W.todense()).ravel(), 1.0 / 8, atol=1e-4)

    def test_regular_tetrahedron_centroid(self):
        keys = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=np.float64)
        W = natinterp3d.get_weights(keys.mean(axis=0, keepdims=True), keys)
        np.testing.assert_allclose(np.asarray(W.todense()).ravel(), 0.25, atol=1e-10)

    def test_octahedron_center(self):
        keys = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]],
                        dtype=np.float64)
        W = natinterp3d.get_weights(np.array([[0.0, 0.0, 0.0]]), keys)
        np.testing.assert_allclose(np.asarray(W.todense()).ravel(), 1.0 / 6, atol=1e-4)


class TestQueryAtDataPoint:  # :184-208
    def test_exact_data_point(self):
        rng = np.random.RandomState(33)
        keys = rng.randn(50, 3)
        for idx in [0, 10, 25, 49]:
            W = natinterp3d.get_weights(keys[idx:idx + 1], keys)
            w = np.asarray(W.todense()).ravel()
            assert w[idx] == pytest.approx(1.0, abs=1e-4)
            np.testing.assert_allclose(np.delete(w, idx), 0.0, atol=1e-4)

    def test_interpolation_at_data_point(self):
        rng = np.random.RandomState(34)
        keys = rng.randn(50, 3)
        values = rng.randn(50)
        for idx in [0, 25, 49]:
            result = natinterp3d.interpolate(keys[idx:idx + 1], keys, values)
            np.testing.assert_allclose(result[0], values[idx], atol=1e-3)


class TestOutsideConvexHull:  # :213-236
    def test_far_away_query(self):
        rng = np.random.RandomState(20)
        keys = rng.randn(100, 3)
        queries = np.array([[100.0, 100.0, 100.0], [-100.0, -100.0, -100.0]])
        W = natinterp3d.get_weights(queries, keys)
        np.testing.assert_allclose(np.asarray(W.sum(axis=1)).ravel(), 0.0, atol=1e-15)

    def test_interpolation_outside(self):
        rng = np.random.RandomState(21)
        keys = rng.randn(100, 3)
        values = rng.randn(100) + 10.0
        result = natinterp3d.interpolate(np.array([[100.0, 100.0, 100.0]]), keys, values)
        assert np.all(np.isfinite(result))


class TestSmallPointSets:  # :239-264
    def test_five_points(self):
        keys = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0.5, 0.5, 0.5]],
                        dtype=np.float64)
        w = np.asarray(natinterp3d.get_weights(np.array([[0.2, 0.2, 0.2]]), keys)
                       .todense()).ravel()
        assert w.sum() == pytest.approx(1.0, abs=1e-8)
        assert np.all(w >= -1e-12)

    def test_four_points(self):
        keys = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
        w = np.asarray(natinterp3d.get_weights(np.array([[0.1, 0.1, 0.1]]), keys)
                       .todense()).ravel()
        assert w.sum() == pytest.approx(1.0, abs=1e-8)
        assert np.all(w >= -1e-12)


class TestSingleQuery:  # :267-276
    def test_single_query(self):
        rng = np.random.RandomState(77)
        keys = rng.randn(50, 3)
        queries = np.array([[0.0, 0.0, 0.0]])
        assert natinterp3d.get_weights(queries, keys).shape == (1, 50)
        result = natinterp3d.interpolate(queries, keys, rng.randn(50))
        assert result.shape == (1,)


class TestLargeCoordinates:  # :279-311
    @pytest.mark.parametrize("seed,transform", [
        (88, lambda k: k + np.array([1e6, -2e6, 3e6])),
        (89, lambda k: k * 1e4),
        (90, lambda k: k * 1e-4),
    ])
    def test_scale(self, seed, transform):
        rng = np.random.RandomState(seed)
        keys = transform(rng.randn(200, 3))
        queries = random_interior_queries(keys, 30, rng)
        W = natinterp3d.get_weights(queries, keys)
        sums = np.asarray(W.sum(axis=1)).ravel()
        interior = sums > 0.5
        assert interior.sum() > 15
        np.testing.assert_allclose(sums[interior], 1.0, atol=1e-6)


class TestOutputShapes:  # :316-344
    @pytest.fixture()
    def data(self):
        rng = np.random.RandomState(1)
        return rng.randn(100, 3), rng.randn(50, 3) * 0.3

    def test_1d_values(self, data):
        keys, queries = data
        assert natinterp3d.interpolate(queries, keys, np.ones(100)).shape == (50,)

    def test_2d_values(self, data):
        keys, queries = data
        assert natinterp3d.interpolate(queries, keys, np.ones((100, 5))).shape == (50, 5)

    def test_weight_matrix_shape(self, data):
        keys, queries = data
        assert natinterp3d.get_weights(queries, keys).shape == (50, 100)

    def test_weight_matrix_is_csr(self, data):
        keys, queries = data
        assert isinstance(natinterp3d.get_weights(queries, keys), scipy.sparse.csr_matrix)


class TestParallelSerialConsistency:  # :347-373
    def test_weights_match(self):
        rng = np.random.RandomState(2)
        keys = rng.randn(200, 3)
        queries = rng.randn(80, 3) * 0.3
        W_par = natinterp3d.get_weights(queries, keys, parallel=True)
        W_ser = natinterp3d.get_weights(queries, keys, parallel=False)
        np.testing.assert_allclose(W_par.toarray(), W_ser.toarray(), atol=1e-12)

    def test_interpolation_match(self):
        rng = np.random.RandomState(3)
        keys = rng.randn(200, 3)
        queries = rng.randn(80, 3) * 0.3
        values = rng.randn(200)
        r_par = natinterp3d.interpolate(queries, keys, values, parallel=True)
        r_ser = natinterp3d.interpolate(queries, keys, values, parallel=False)
        np.testing.assert_allclose(r_par, r_ser, atol=1e-12)

    def test_explicit_thread_count(self):
        rng = np.random.RandomState(4)
        keys = rng.randn(100, 3)
        queries = rng.randn(30, 3) * 0.3
        W1 = natinterp3d.get_weights(queries, keys, parallel=True, num_threads=1)
        W2 = natinterp3d.get_weights(queries, keys, parallel=True, num_threads=4)
        W_ser = natinterp3d.get_weights(queries, keys, parallel=False)
        np.testing.assert_allclose(W1.toarray(), W_ser.toarray(), atol=1e-12)
        np.testing.assert_allclose(W2.toarray(), W_ser.toarray(), atol=1e-12)


class TestInterpolatorReuse:  # :376-399
    def test_multiple_get_weights(self):
        rng = np.random.RandomState(5)
        keys = rng.randn(100, 3)
        interp = natinterp3d.Interpolator(keys)
        q1 = rng.randn(20, 3) * 0.3
        q2 = rng.randn(30, 3) * 0.3
        W1a = interp.get_weights(q1)
        W2 = interp.get_weights(q2)
        W1b = interp.get_weights(q1)
        np.testing.assert_allclose(W1a.toarray(), W1b.toarray(), atol=1e-15)
        assert W2.shape == (30, 100)

    def test_multiple_interpolate(self):
        rng = np.random.RandomState(6)
        keys = rng.randn(100, 3)
        values = rng.randn(100)
        interp = natinterp3d.Interpolator(keys)
        queries = rng.randn(20, 3) * 0.3
        np.testing.assert_allclose(interp.interpolate(queries, values),
                                   interp.interpolate(queries, values), atol=1e-15)


class TestDtypeHandling:  # :402-421
    def test_float32_input(self):
        rng = np.random.RandomState(50)
        keys = rng.randn(100, 3).astype(np.float32)
        queries = (rng.randn(20, 3) * 0.3).astype(np.float32)
        values = rng.randn(100).astype(np.float32)
        result = natinterp3d.interpolate(queries, keys, values)
        assert result.dtype == np.float64 or result.dtype == np.float32
        assert result.shape == (20,)

    def test_float32_weights(self):
        rng = np.random.RandomState(51)
        keys = rng.randn(100, 3).astype(np.float32)
        queries = (rng.randn(20, 3) * 0.3).astype(np.float32)
        W = natinterp3d.get_weights(queries, keys)
        assert W.shape == (20, 100)
        sums = np.asarray(W.sum(axis=1)).ravel()
        np.testing.assert_allclose(sums[sums > 0.5], 1.0, atol=1e-6)


class TestConvenienceFunctions:  # :424-444
    def test_interpolate_matches_class(self):
        rng = np.random.RandomState(60)
        keys = rng.randn(100, 3)
        queries = rng.randn(20, 3) * 0.3
        values = rng.randn(100)
        r1 = natinterp3d.interpolate(queries, keys, values, parallel=False)
        r2 = natinterp3d.Interpolator(keys, parallel=False).interpolate(queries, values)
        np.testing.assert_allclose(r1, r2, atol=1e-15)

    def test_get_weights_matches_class(self):
        rng = np.random.RandomState(61)
        keys = rng.randn(100, 3)
        queries = rng.randn(20, 3) * 0.3
        W1 = natinterp3d.get_weights(queries, keys, parallel=False)
        W2 = natinterp3d.Interpolator(keys, parallel=False).get_weights(queries)
        np.testing.assert_allclose(W1.toarray(), W2.toarray(), atol=1e-15)


class TestDeterminism:  # :449-466
    def test_weights_deterministic(self):
        rng = np.random.RandomState(70)
        keys = rng.randn(200, 3)
        queries = rng.randn(50, 3) * 0.3
        W1 = natinterp3d.get_weights(queries, keys, parallel=True)
        W2 = natinterp3d.get_weights(queries, keys, parallel=True)
        np.testing.assert_array_equal(W1.toarray(), W2.toarray())

    def test_serial_deterministic(self):
        rng = np.random.RandomState(71)
        keys = rng.randn(200, 3)
        queries = rng.randn(50, 3) * 0.3
        W1 = natinterp3d.get_weights(queries, keys, parallel=False)
        W2 = natinterp3d.get_weights(queries, keys, parallel=False)
        np.testing.assert_array_equal(W1.toarray(), W2.toarray())


class TestSparsity:  # :471-484
    def test_few_nonzeros_per_row(self):
        rng = np.random.RandomState(80)
        keys = rng.randn(1000, 3)
        queries = rng.randn(50, 3) * 0.3
        W = natinterp3d.get_weights(queries, keys)
        for i in range(W.shape[0]):
            if W[i].nnz > 0:
                assert W[i].nnz < 100


class TestOriginal:  # :489-503
    def test_natinterp3d(self):
        rng = np.random.RandomState(10)
        keys = rng.randn(10000, 3)

        def func(x):
            x = x * 0.1
            return np.sin(x[:, 0]) + np.cos(x[:, 1]) + np.tan(x[:, 2])

        values = func(keys)
        queries = rng.randn(100, 3)
        interp_values = natinterp3d.interpolate(queries, keys, values)
        assert np.mean(np.abs(interp_values - func(queries))) < 1e-3


class TestAgainstReferenceFixtures:
    """Every golden fixture (the compiled reference's CSR on the inputs of its
    own tests) reproduced bit for bit through the public API."""

    @pytest.mark.parametrize("name", sorted(__import__("tests.golden.cases",
                                                       fromlist=["CASES"]).CASES))
    def test_fixture(self, name):
        import os
        small = np.load(os.path.join(os.path.dirname(__file__), "golden", "small_cases.npz"))
        keys = small[f"{name}__keys"]
        q = small[f"{name}__queries"]
        W = natinterp3d.get_weights(q, keys)
        assert np.array_equal(W.indptr, small[f"{name}__indptr"])
        assert np.array_equal(W.indices, small[f"{name}__indices"])
        d, e = W.data, small[f"{name}__data"]
        assert ((d == e) | (np.isnan(d) & np.isnan(e))).all()
