"""nat_weights_noalloc -- the entry point that allocates nothing -- against
nat_weights_compute + nat_weights_emit through the C ABI: bit-identical CSR,
two calls in a row with the second one smaller (the shared scratch keeps the
first call's contents), on a fresh mesh (no step-3 table yet), on rows that
need the wave-per-query level, in mixed precision, and its size checks."""
import ctypes

import numpy as np
import pytest

from natinterp3d_amd import _lib, workloads
from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP

pytestmark = pytest.mark.gpu

MAX_QUERIES = 32768


def say(msg):
    print(msg, flush=True)


def noalloc(m, qd, capacity):
    import torch
    M = qd.shape[0]
    val = torch.full((max(capacity, 1),), -1.0, dtype=torch.float64, device=qd.device)
    col = torch.full((max(capacity, 1),), -1, dtype=torch.int32, device=qd.device)
    crow = torch.full((M + 1,), -1, dtype=torch.int32, device=qd.device)
    nnz = ctypes.c_int64(-5)
    stream = ctypes.c_void_p(torch.cuda.current_stream(qd.device).cuda_stream)
    rc = _lib.lib().nat_weights_noalloc(
        m.handle, ctypes.c_void_p(qd.data_ptr() if M else 0), M,
        ctypes.c_void_p(val.data_ptr()), ctypes.c_void_p(col.data_ptr()),
        ctypes.c_void_p(crow.data_ptr()), capacity, stream, ctypes.byref(nnz))
    torch.cuda.synchronize()
    return rc, int(nnz.value), crow, col[:max(int(nnz.value), 0)], val[:max(int(nnz.value), 0)]


def check_equal(m, qd, label, per_row=64):
    import torch
    rc, nnz, crow, col, val = noalloc(m, qd, qd.shape[0] * per_row)
    assert rc == 0, f"{label}: rc {rc}: {_lib.last_error()}"
    c2, k2, v2 = m.weights_device(qd)
    torch.cuda.synchronize()
    assert nnz == k2.numel(), f"{label}: nnz {nnz} vs {k2.numel()}"
    assert torch.equal(crow, c2), f"{label}: row pointers differ"
    assert torch.equal(col, k2), f"{label}: columns differ"
    assert torch.equal(val, v2), f"{label}: weights differ"
    say(f"[noalloc {label}] M={qd.shape[0]} nnz={nnz}: equal to compute + emit")
    return nnz


def centred_mesh(keys):
    keys = np.ascontiguousarray(keys, np.float64)
    c = keys.mean(axis=0)
    return MeshAndVerticesHIP(np.ascontiguousarray(keys - c), device=0), c


def test_noalloc_equals_compute_emit_two_calls():
    import torch
    k, q, _ = workloads.plain_at_scale(n_keys=20_000, n_queries=MAX_QUERIES)
    m, c = centred_mesh(k)
    q1 = torch.tensor(q - c, device="cuda")
    q2 = torch.tensor(q[:9000][::-1].copy() - c, device="cuda")
    # first on the fresh mesh: no step-3 table yet, so the geometry runs
    # step 3 inline; the compute + emit that follows builds the table
    check_equal(m, q1, "fresh mesh, full size")
    check_equal(m, q2, "second call, smaller")
    check_equal(m, q1, "third call, with the step-3 table")
    m.set_policy(geom_lanes=1)
    check_equal(m, q2, "one geometry lane per row")
    m.set_policy(geom_lanes=2)
    m.set_precision("mixed")
    check_equal(m, q1, "mixed precision")
    check_equal(m, q2[:100], "mixed precision, 100 queries")


def test_noalloc_wave_rows():
    """Queries inside a hollow shell: every row leaves the group kernel for
    the wave-per-query level (cavities of hundreds of tets)."""
    import torch
    k, q, _, _ = workloads.shell(n_keys=3000, n_queries=1500)
    m, c = centred_mesh(k)
    qd = torch.tensor(q - c, device="cuda")
    m.set_timing(True)
    m.weights_device(qd)
    assert m.timings()["wave_items"] > 1000
    check_equal(m, qd, "shell rows", per_row=1000)
    check_equal(m, qd[:300], "shell rows, smaller", per_row=1000)


def test_noalloc_limits():
    import torch
    k, q, _ = workloads.uniform_small(n_keys=2000, n_queries=MAX_QUERIES + 1)
    m, c = centred_mesh(k)
    qd = torch.tensor(q - c, device="cuda")
    rc, _, _, _, _ = noalloc(m, qd, 10)
    assert rc == -1 and "at most" in _lib.last_error()
    # output capacity too small: -7, the needed size reported, nothing written
    rc, nnz, crow, _, _ = noalloc(m, qd[:1000], 100)
    assert rc == -7 and nnz > 100
    assert (crow == -1).all().item()
    rc, nnz2, crow, col, val = noalloc(m, qd[:1000], nnz)
    assert rc == 0 and nnz2 == nnz and crow[-1].item() == nnz
    # no queries: rowptr[0] = 0
    rc, nnz, crow, _, _ = noalloc(m, qd[:0], 0)
    assert rc == 0 and nnz == 0 and crow[0].item() == 0
