"""The boundary's overflow convention (natural_insertionfree.c:805-812,
natinterp3d_cython.pyx:84-88): a total above INT32_MAX entries makes
nat_get_weights return -1 with nothing allocated, and the plugin object
raises OverflowError; the device-resident calls follow the same rule.

2^31 real entries would need ~26 GB of output, so the library's test hook
NATINTERP3D_TEST_FORGE_NNZ adds entries to the total the checks see; the rows
are not changed by it, and a total of exactly INT32_MAX still passes (the
reference's test is a strict >).
"""
import ctypes

import numpy as np
import pytest

from natinterp3d_amd import _lib, workloads
from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP

pytestmark = pytest.mark.gpu

INT32_MAX = 2**31 - 1


@pytest.fixture(scope="module")
def case():
    k, q, _ = workloads.uniform_small(n_keys=2000, n_queries=5000)
    c = k.mean(axis=0)
    m = MeshAndVerticesHIP(np.ascontiguousarray(k - c), device=0)
    qc = np.ascontiguousarray(q - c)
    return m, qc, m.get_weights(qc)


def test_get_weights_overflow_raises_and_allocates_nothing(case, monkeypatch):
    m, qc, W = case
    monkeypatch.setenv("NATINTERP3D_TEST_FORGE_NNZ", str(INT32_MAX - W.nnz + 1))
    with pytest.raises(OverflowError, match="int32"):
        m.get_weights(qc)
    # the raw entry point: -1, the callee-allocated outputs left NULL
    L = _lib.lib()
    rowptr = np.zeros(len(qc) + 1, np.int32)
    vals = ctypes.POINTER(ctypes.c_double)()
    cols = ctypes.POINTER(ctypes.c_int32)()
    rc = L.nat_get_weights(qc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(qc),
                           m.handle, m.numDataPoints, ctypes.byref(vals), ctypes.byref(cols),
                           rowptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    assert rc == -1
    assert not vals and not cols


def test_total_of_exactly_int32_max_passes(case, monkeypatch):
    m, qc, W = case
    monkeypatch.setenv("NATINTERP3D_TEST_FORGE_NNZ", str(INT32_MAX - W.nnz))
    W2 = m.get_weights(qc)
    assert np.array_equal(W2.indptr, W.indptr) and np.array_equal(W2.indices, W.indices)
    assert np.array_equal(W2.data, W.data)


def test_device_calls_overflow(case, monkeypatch):
    import torch
    m, qc, W = case
    qd = torch.tensor(qc, device="cuda")
    monkeypatch.setenv("NATINTERP3D_TEST_FORGE_NNZ", str(INT32_MAX))
    with pytest.raises(OverflowError):
        m.weights_device(qd)
    with pytest.raises(OverflowError):
        m.compute_device(qd)
    # nat_weights_emit itself returns -1 after such a compute, before writing
    L = _lib.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    nnz = ctypes.c_int64()
    assert L.nat_weights_compute(m.handle, ctypes.c_void_p(qd.data_ptr()), len(qc), stream,
                                 ctypes.byref(nnz)) == 0
    assert nnz.value == W.nnz + INT32_MAX
    crow = torch.full((len(qc) + 1,), -7, dtype=torch.int32, device="cuda")
    val = torch.empty(1, dtype=torch.float64, device="cuda")
    col = torch.empty(1, dtype=torch.int32, device="cuda")
    rc = L.nat_weights_emit(m.handle, ctypes.c_void_p(val.data_ptr()),
                            ctypes.c_void_p(col.data_ptr()), ctypes.c_void_p(crow.data_ptr()),
                            stream)
    torch.cuda.synchronize()
    assert rc == -1
    assert (crow == -7).all().item(), "emit wrote the row pointers despite the overflow"
    # without the hook the same mesh computes again
    monkeypatch.delenv("NATINTERP3D_TEST_FORGE_NNZ")
    crow, col, val = m.weights_device(qd)
    assert np.array_equal(crow.cpu().numpy(), W.indptr)
