"""Batched problems (BatchInterpolator): every problem's CSR, product and
gradient must equal those of the problem run alone -- bit for bit for the
weights and products, to 1e-12 for the gradients (atomic order)."""
import numpy as np
import pytest

import natinterp3d_amd
from natinterp3d_amd import workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def make_problems(seed, n_problems=10):
    rng = np.random.default_rng(seed)
    keys, queries = [], []
    for p in range(n_problems):
        nk = int(rng.integers(20, 4000))
        nq = int(rng.integers(0, 2500)) if p != 3 else 0  # problem 3: no queries
        side = rng.uniform(0.5, 2.0, 3)
        keys.append(rng.random((nk, 3)) * side)
        queries.append(rng.random((nq, 3)) * side)
    # a hollow shell: its rows come from the larger-capacity levels
    k, q, _, _ = workloads.shell(seed=5, n_keys=1500, n_queries=200)
    keys.append(k)
    queries.append(q)
    return keys, queries


def assert_csr_equal(W, Wo, what):
    assert W.shape == Wo.shape, what
    assert np.array_equal(W.indptr, Wo.indptr), f"{what}: indptr differs"
    assert np.array_equal(W.indices, Wo.indices), f"{what}: indices differ"
    assert np.array_equal(W.data, Wo.data), f"{what}: data differs"


def test_batch_weights_equal_single_problems(oracle_mod):
    keys, queries = make_problems(21)
    say(f"[batch] {len(keys)} problems, {sum(len(q) for q in queries)} queries")
    b = natinterp3d_amd.BatchInterpolator(keys)
    assert b.info()["n_problems"] == len(keys)
    Ws = b.get_weights(queries)
    for p, (k, q) in enumerate(zip(keys, queries)):
        single = natinterp3d_amd.Interpolator(k)
        Wo = oracle_mod.OracleMesh(single.mesh_and_vertices.export()).get_weights(
            np.ascontiguousarray(q, np.float64) - single._centroid)
        assert_csr_equal(Ws[p], Wo, f"problem {p}")
    say("[batch] weights equal the oracle per problem")


def test_batch_interpolate_equals_single():
    keys, queries = make_problems(22, n_problems=6)
    rng = np.random.default_rng(23)
    for qdt, vdt, shape in ((np.float64, np.float64, (4,)), (np.float32, np.float32, ()),
                            (np.float64, np.float32, (3,))):
        values = [rng.standard_normal((len(k),) + shape).astype(vdt) for k in keys]
        qs = [q.astype(qdt) for q in queries]
        b = natinterp3d_amd.BatchInterpolator(keys)
        ys = b.interpolate(qs, values)
        for p in range(len(keys)):
            ref = natinterp3d_amd.Interpolator(keys[p]).interpolate(qs[p], values[p])
            assert ys[p].dtype == ref.dtype and ys[p].shape == ref.shape
            assert np.array_equal(ys[p], ref), f"problem {p} ({qdt.__name__}, {vdt.__name__})"
    say("[batch] products equal the single-problem products")


def test_batch_backward_equals_single():
    import torch
    keys, queries = make_problems(24, n_problems=5)
    rng = np.random.default_rng(25)
    values = [rng.standard_normal((len(k), 2)) for k in keys]
    b = natinterp3d_amd.BatchInterpolator(keys)
    off = np.zeros(len(keys) + 1, np.int64)
    off[1:] = np.cumsum([len(q) for q in queries])
    dev = torch.device("cuda", 0)
    qt = torch.tensor(np.concatenate(queries), device=dev, requires_grad=True)
    vt = torch.tensor(np.concatenate(values), device=dev, requires_grad=True)
    g = torch.tensor(rng.standard_normal((int(off[-1]), 2)), device=dev)
    y = b.interpolate_torch(qt, off, vt)
    (y * g).sum().backward()
    voff = b.vertex_offsets
    for p in range(len(keys)):
        qp = torch.tensor(queries[p], device=dev, requires_grad=True)
        vp = torch.tensor(values[p], device=dev, requires_grad=True)
        yp = natinterp3d_amd.Interpolator(keys[p]).interpolate_torch(qp, vp)
        (yp * g[off[p]:off[p + 1]]).sum().backward()
        assert torch.equal(y[off[p]:off[p + 1]], yp.detach()), f"problem {p}: forward"
        torch.testing.assert_close(qt.grad[off[p]:off[p + 1]], qp.grad, rtol=1e-12, atol=1e-12)
        torch.testing.assert_close(vt.grad[voff[p]:voff[p + 1]], vp.grad, rtol=1e-12,
                                   atol=1e-12)
    say("[batch] gradients equal the single-problem gradients")
