"""The query order of a compute is the stable sort of the queries' Morton
keys (problem-major, then the 3*bits-bit Morton code of the query's cell in
the data bounding box, equal keys in index order): a function of the queries
alone.  Rows never depend on it, but which kernel makes a row does (queries
are searched in quads of consecutive queries), so the same queries must give
the same mixed-precision weights on every call."""
import ctypes

import numpy as np
import pytest

from natinterp3d_amd import _lib, workloads
from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP

pytestmark = pytest.mark.gpu


def spread3(x):
    x = x.astype(np.int64) & 0x3ff
    x = (x | (x << 16)) & 0x030000ff
    x = (x | (x << 8)) & 0x0300f00f
    x = (x | (x << 4)) & 0x030c30c3
    x = (x | (x << 2)) & 0x09249249
    return x


def expected_order(kc, qc):
    """Stable argsort of the keys, restating the key arithmetic in float64."""
    M = len(qc)
    bits = 1
    while bits < 7 and (1 << (3 * (bits + 1))) <= 2 * M and (1 << (3 * (bits + 1))) <= (1 << 24):
        bits += 1
    lo, hi = kc.min(axis=0), kc.max(axis=0)
    cells = []
    for d in range(3):
        r = hi[d] - lo[d]
        s = float(1 << bits) / r if r > 0 else 1.0
        c = (qc[:, d] - lo[d]) * s
        k = np.where((c >= 0.0) & (c < 2147483647.0), np.floor(np.nan_to_num(c, nan=0.0)), 0.0)
        k = np.where(c >= 2147483647.0, (1 << bits) - 1, k)
        k = np.minimum(k, (1 << bits) - 1)
        cells.append(k.astype(np.int64))
    key = spread3(cells[0]) | (spread3(cells[1]) << 1) | (spread3(cells[2]) << 2)
    return np.argsort(key, kind="stable").astype(np.int32)


@pytest.mark.parametrize("n_queries", [1, 999, 70_000, 300_001])
def test_query_order_is_the_stable_morton_sort(n_queries):
    import torch
    k, q, _ = workloads.plain_at_scale(n_keys=5000, n_queries=n_queries)
    c = k.mean(axis=0)
    kc = np.ascontiguousarray(k - c)
    qc = np.ascontiguousarray(q - c)
    qc[: min(10, len(qc))] = qc[0]  # equal keys keep index order
    m = MeshAndVerticesHIP(kc, device=0)
    qd = torch.tensor(qc, device="cuda")
    m.compute_device(qd)
    # the saved topology holds the compute's query order at bytes 64 + 8M ..
    L = _lib.lib()
    nb = ctypes.c_int64()
    assert L.nat_topology_save(m.handle, None, 0, ctypes.byref(nb), None) == 0
    topo = torch.empty(int(nb.value), dtype=torch.uint8, device="cuda")
    assert L.nat_topology_save(m.handle, ctypes.c_void_p(topo.data_ptr()), topo.numel(),
                               ctypes.byref(nb), None) == 0
    M = len(qc)
    got = topo[64 + 8 * M: 64 + 12 * M].cpu().numpy().view(np.int32)
    assert np.array_equal(got, expected_order(kc, qc))


def test_mixed_rows_are_reproducible():
    import torch
    k, q, _ = workloads.plain_at_scale(n_keys=20_000, n_queries=100_000)
    c = k.mean(axis=0)
    m = MeshAndVerticesHIP(np.ascontiguousarray(k - c), device=0)
    qd = torch.tensor(q - c, device="cuda")
    m.set_precision("mixed")
    a = [t.clone() for t in m.weights_device(qd)]
    for _ in range(3):
        b = m.weights_device(qd)
        assert all(torch.equal(x, y) for x, y in zip(a, b)), "mixed rows changed between calls"
