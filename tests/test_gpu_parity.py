"""GPU parity: the HIP query path against the oracle restatement (and the
compiled reference where oracle/_ref is present), bit for bit.

CSR indptr, indices (column order included) and data must be identical;
rows are computed in FP64 with the reference's operation order, so the
tolerance is zero.
"""
import numpy as np
import pytest

import natinterp3d_amd
from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def assert_csr_equal(W, Wo, what):
    assert W.shape == Wo.shape, what
    assert np.array_equal(W.indptr, Wo.indptr), f"{what}: indptr differs"
    assert np.array_equal(W.indices, Wo.indices), f"{what}: indices differ"
    same = (W.data == Wo.data) | (np.isnan(W.data) & np.isnan(Wo.data))
    assert same.all(), f"{what}: {np.count_nonzero(~same)} weights differ"


def run_case(oracle_mod, keys, queries, label, ref=None):
    keys = np.ascontiguousarray(keys, np.float64)
    c = keys.mean(axis=0)
    kc = keys - c
    qc = np.ascontiguousarray(queries, np.float64) - c
    say(f"[case {label}] N={len(keys)} M={len(queries)} build")
    m = MeshAndVerticesHIP(kc, device=0)
    say(f"[case {label}] gpu get_weights")
    W = m.get_weights(qc)
    Wo = oracle_mod.OracleMesh(m.export()).get_weights(qc)
    assert_csr_equal(W, Wo, f"{label} vs oracle")
    if ref is not None:
        r = ref.RefMesh(kc)
        Wr = r.get_weights(qc, num_threads=2)
        assert_csr_equal(W, Wr, f"{label} vs reference")
    return m, W


SMALL = {
    "cube": (np.array([[-1, -1, -1], [-1, -1, 1], [-1, 1, -1], [-1, 1, 1],
                       [1, -1, -1], [1, -1, 1], [1, 1, -1], [1, 1, 1]], float),
             np.array([[0.0, 0, 0], [0.3, -0.2, 0.1]])),
    "octahedron": (np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1],
                             [0, 0, -1]], float), np.array([[0.0, 0, 0]])),
    "regular_tet": (np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float),
                    np.array([[0.0, 0, 0]])),
    "four": (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float),
             np.array([[0.1, 0.1, 0.1]])),
    "five": (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0.5, 0.5, 0.5]], float),
             np.array([[0.2, 0.2, 0.2]])),
}


@pytest.mark.parametrize("name", sorted(SMALL))
def test_small_known_sets(oracle_mod, name):
    keys, q = SMALL[name]
    run_case(oracle_mod, keys, q, name)


def test_edge_queries(oracle_mod):
    """Data-point hits (coincident path), far queries (outside the super
    simplex: empty rows), extrapolation inside the super simplex, NaN and inf
    rows, duplicate queries."""
    rng = np.random.RandomState(33)
    keys = rng.randn(50, 3)
    q = np.vstack([keys[[0, 10, 25, 49]],
                   [[100.0, 100, 100], [-100.0, -100, -100], [3.0, 3, 3], [-2.5, 0, 1]],
                   [[np.nan, 0, 0], [np.inf, 0, 0], [0, 0, 0], [0, 0, 0]]])
    m, W = run_case(oracle_mod, keys, q, "edge_queries")
    for i, idx in enumerate([0, 10, 25, 49]):
        row = W[i]
        assert row.nnz == 1 and row.indices[0] == idx and row.data[0] == 1.0
    assert W[4].nnz == 0 and W[5].nnz == 0
    assert W[8].nnz == 0 and W[9].nnz == 0


def test_empty_query_set(oracle_mod):
    rng = np.random.RandomState(1)
    keys = rng.randn(100, 3)
    m = MeshAndVerticesHIP(np.ascontiguousarray(keys - keys.mean(0)), device=0)
    W = m.get_weights(np.zeros((0, 3)))
    assert W.shape == (0, 100) and W.nnz == 0


def test_grid_with_perturbations(oracle_mod):
    xs = np.linspace(-1, 1, 6)
    keys = np.array(np.meshgrid(xs, xs, xs)).reshape(3, -1).T
    q = np.random.RandomState(7).uniform(-0.8, 0.8, (300, 3))
    run_case(oracle_mod, keys, q, "grid6")


def test_config1_uniform_10k_100k(oracle_mod):
    """10,000 keys uniform in [0,1]^3, 100,000 queries in [0.05,0.95]^3."""
    rng = np.random.default_rng(0)
    keys = rng.random((10_000, 3))
    q = 0.05 + 0.9 * rng.random((100_000, 3))
    m, W = run_case(oracle_mod, keys, q, "config1")
    assert W.nnz > 15 * 100_000


def test_large_cavities_escalate(oracle_mod):
    """Keys on a sphere, queries inside: cavities of hundreds of tets exceed
    the first capacity level and must be finished by the larger ones."""
    rng = np.random.default_rng(5)
    d = rng.standard_normal((3000, 3))
    keys = d / np.linalg.norm(d, axis=1, keepdims=True)
    q = 0.3 * rng.standard_normal((2000, 3)).clip(-2, 2)
    m, W = run_case(oracle_mod, keys, q, "sphere_shell")
    assert np.diff(W.indptr).max() > 96


def test_repeated_calls_identical(oracle_mod):
    rng = np.random.default_rng(11)
    keys = rng.random((3000, 3))
    q1 = rng.random((20_000, 3))
    q2 = rng.random((7_000, 3))
    c = keys.mean(0)
    m = MeshAndVerticesHIP(np.ascontiguousarray(keys - c), device=0)
    a = m.get_weights(q1 - c)
    m.get_weights(q2 - c)
    b = m.get_weights(q1 - c)
    assert_csr_equal(a, b, "repeat")


def test_device_path_matches_host_path(oracle_mod):
    import torch
    rng = np.random.default_rng(12)
    keys = rng.random((4000, 3))
    q = rng.random((10_000, 3))
    c = keys.mean(0)
    m = MeshAndVerticesHIP(np.ascontiguousarray(keys - c), device=0)
    W = m.get_weights(q - c)
    crow, col, val = m.weights_device(torch.from_numpy(q - c).cuda())
    torch.cuda.synchronize()
    assert np.array_equal(crow.cpu().numpy(), W.indptr)
    assert np.array_equal(col.cpu().numpy(), W.indices)
    assert np.array_equal(val.cpu().numpy(), W.data)


@pytest.mark.parametrize("vdt,qdt", [(np.float64, np.float64), (np.float32, np.float64),
                                     (np.float64, np.float32), (np.float32, np.float32)])
def test_interpolate_matches_scipy_bitwise(oracle_mod, vdt, qdt):
    rng = np.random.default_rng(13)
    keys = rng.random((3000, 3))
    q = (0.1 + 0.8 * rng.random((5000, 3))).astype(qdt)
    vals = rng.standard_normal((3000, 3)).astype(vdt)
    it = natinterp3d_amd.Interpolator(keys)
    y = it.interpolate(q, vals)
    W = it.get_weights(q)
    expect = W @ vals
    assert y.dtype == expect.dtype
    assert np.array_equal(y, expect)
    y1 = it.interpolate(q, vals[:, 0])
    assert y1.shape == (5000,) and np.array_equal(y1, W @ vals[:, 0])


@pytest.mark.parametrize("vdt", [np.float64, np.float32])
def test_spmm_widths_and_row_order_bitwise(vdt):
    """W @ values for value widths 1..7 (the 16-lane kernel up to 4, one
    thread per (row, column) beyond) equals scipy's product bit for bit,
    with rows visited in index order and in the compute's spatial order;
    rows longer than 16 entries cross the lane chunk."""
    import torch
    from natinterp3d_amd.natinterp3d import row_order, spmm_device
    rng = np.random.default_rng(17)
    keys = rng.random((2500, 3))
    q = 0.05 + 0.9 * rng.random((4000, 3))
    it = natinterp3d_amd.Interpolator(keys)
    m = it.mesh_and_vertices
    W = it.get_weights(q)
    assert np.diff(W.indptr).max() > 16
    qd = torch.from_numpy(q - it._centroid).cuda()
    crow, col, val = m.weights_device(qd)
    order = row_order(m.handle, len(q))
    for D in range(1, 8):
        vals = rng.standard_normal((2500, D)).astype(vdt)
        expect = W @ vals
        xd = torch.from_numpy(vals).cuda()
        for o in (None, order):
            y = spmm_device(crow, col, val, xd, len(q), order=o).cpu().numpy()
            assert y.dtype == expect.dtype
            assert np.array_equal(y, expect), (D, o is not None)


def test_host_call_large_transfers_match_device_csr():
    """get_weights on host buffers large enough for the pinned multi-lane
    copies (9.6 MB of queries in, ~75 MB of CSR out, slices not a multiple of
    the chunk) returns the device-resident CSR byte for byte, twice in a row
    (the second call reuses the staging)."""
    import torch
    rng = np.random.default_rng(23)
    keys = rng.random((20000, 3))
    q = 0.02 + 0.96 * rng.random((400003, 3))
    it = natinterp3d_amd.Interpolator(keys)
    crow, col, val = it.mesh_and_vertices.weights_device(
        torch.from_numpy(q - it._centroid).cuda())
    torch.cuda.synchronize()
    crow, col, val = crow.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy()
    assert val.nbytes > 8 << 20 and 3 * 8 * len(q) > 8 << 20
    for _ in range(2):
        W = it.get_weights(q)
        assert np.array_equal(W.indptr, crow)
        assert np.array_equal(W.indices, col)
        assert np.array_equal(W.data, val)


@pytest.mark.parametrize("vdt", [np.float64, np.float32])
def test_pool_product_equals_csr_product(vdt):
    """W @ values read from the row pool (nat_weights_spmm) equals the product
    on the emitted CSR bit for bit, value widths 1..4, single mesh and batch;
    a width of 5 and a query count other than the last compute's are refused."""
    import torch
    from natinterp3d_amd.natinterp3d import spmm_device, weights_spmm
    rng = np.random.default_rng(29)
    keys = rng.random((2500, 3))
    q = 0.05 + 0.9 * rng.random((4000, 3))
    it = natinterp3d_amd.Interpolator(keys)
    m = it.mesh_and_vertices
    crow, col, val = m.weights_device(torch.from_numpy(q - it._centroid).cuda())
    for D in range(1, 5):
        xd = torch.from_numpy(rng.standard_normal((2500, D)).astype(vdt)).cuda()
        assert torch.equal(weights_spmm(m.handle, len(q), xd),
                           spmm_device(crow, col, val, xd, len(q))), D
    with pytest.raises(RuntimeError, match="1 to 4"):
        weights_spmm(m.handle, len(q), torch.zeros((2500, 5), dtype=torch.float64).cuda())
    with pytest.raises(RuntimeError, match="queries"):
        weights_spmm(m.handle, len(q) + 1, xd)
    with pytest.raises(RuntimeError, match="data points"):
        weights_spmm(m.handle, len(q), xd[:-1])
    ks = [rng.random((n, 3)) for n in (800, 1500, 300)]
    qs = [0.05 + 0.9 * rng.random((n, 3)) for n in (900, 2000, 400)]
    b = natinterp3d_amd.BatchInterpolator(ks)
    qc, off = b._centre(qs)
    crow, col, val = b.weights_device(torch.from_numpy(qc).cuda(), off)
    xd = torch.from_numpy(rng.standard_normal((2600, 3)).astype(vdt)).cuda()
    assert torch.equal(weights_spmm(b.handle, int(off[-1]), xd),
                       spmm_device(crow, col, val, xd, int(off[-1])))
