"""Wave-per-query rows (cavities beyond the group kernel's LDS tables): the
rows must equal the oracle's bit for bit and the thread-per-query level's, the
gradients the thread-per-query level's (1e-12: order of the sums), for
hollow shells, clustered keys, batches and graph replays."""
import numpy as np
import pytest
import torch

import natinterp3d_amd
from natinterp3d_amd import workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def csr_equal(a, b):
    return (np.array_equal(a.indptr, b.indptr) and np.array_equal(a.indices, b.indices)
            and np.array_equal(a.data, b.data))


@pytest.mark.parametrize("n_keys,n_queries,seed", [(1500, 400, 5), (6000, 600, 9)])
def test_shell_rows_wave_equal_oracle_and_thread_level(oracle_mod, n_keys, n_queries, seed):
    k, q, _, _ = workloads.shell(seed=seed, n_keys=n_keys, n_queries=n_queries)
    interp = natinterp3d_amd.Interpolator(k)
    mesh = interp.mesh_and_vertices
    mesh.set_timing(True)
    W = interp.get_weights(q)
    t = mesh.timings()
    say(f"[wave] shell {n_keys} keys: wave rows {t['wave_items']}, levels {t['rows_items']}")
    assert t["wave_items"] > 0.9 * n_queries  # these rows leave the group kernel
    Wo = oracle_mod.OracleMesh(mesh.export()).get_weights(
        np.ascontiguousarray(q, np.float64) - interp._centroid)
    assert csr_equal(W, Wo), "wave rows differ from the oracle"
    mesh.set_policy(wave_rows=False)
    assert csr_equal(interp.get_weights(q), W), "thread-per-query rows differ"
    mesh.set_policy(wave_rows=True, budgets=[0])  # every row to the wave level
    assert csr_equal(interp.get_weights(q), W), "all-wave rows differ"


def test_uniform_rows_all_through_wave_level(oracle_mod):
    k, q, _ = workloads.uniform_small(seed=11, n_keys=4000, n_queries=3000)
    interp = natinterp3d_amd.Interpolator(k)
    mesh = interp.mesh_and_vertices
    ref = interp.get_weights(q)
    mesh.set_policy(budgets=[0])
    mesh.set_timing(True)
    W = interp.get_weights(q)
    t = mesh.timings()
    assert t["wave_items"] == len(q)
    assert csr_equal(W, ref)
    Wo = oracle_mod.OracleMesh(mesh.export()).get_weights(
        np.ascontiguousarray(q, np.float64) - interp._centroid)
    assert csr_equal(W, Wo)


def test_wave_gradients_match_thread_level():
    k, q, v, _ = workloads.shell(seed=12, n_keys=2000, n_queries=500)
    interp = natinterp3d_amd.Interpolator(k)
    mesh = interp.mesh_and_vertices
    dev = torch.device("cuda", 0)
    qt = torch.tensor(q, device=dev, requires_grad=True)
    vt = torch.tensor(v, device=dev, requires_grad=True)
    g = torch.randn(len(q), 3, dtype=torch.float64, device=dev,
                    generator=torch.Generator(device=dev).manual_seed(3))
    mesh.set_timing(True)
    y = interp.interpolate_torch(qt, vt)
    dq, dv = torch.autograd.grad(y, (qt, vt), g)
    t = mesh.timings()
    say(f"[wave] grad: wave rows {t['grad_wave_items']}, levels {t['grad_items']}")
    assert t["grad_wave_items"] > 0.9 * len(q)
    mesh.set_policy(wave_rows=False)
    y1 = interp.interpolate_torch(qt, vt)
    dq1, dv1 = torch.autograd.grad(y1, (qt, vt), g)
    assert torch.equal(y, y1)
    torch.testing.assert_close(dq, dq1, rtol=1e-11, atol=1e-12)
    torch.testing.assert_close(dv, dv1, rtol=1e-11, atol=1e-12)
    # finite differences on a few rows (Richardson, as in the gradient tests)
    sel = torch.arange(0, len(q), 25, device=dev)
    q0 = qt.detach()[sel]
    mesh.set_policy(wave_rows=True)

    def f(qq):
        with torch.no_grad():
            return (interp.interpolate_torch(qq, vt.detach()) * g[sel]).sum(dim=1)

    fd = torch.empty_like(q0)
    for kk in range(3):
        e = torch.zeros_like(q0)
        h = 1e-5
        e[:, kk] = h
        d1 = (f(q0 + e) - f(q0 - e)) / (2 * h)
        e[:, kk] = h / 2
        d2 = (f(q0 + e) - f(q0 - e)) / h
        fd[:, kk] = (4 * d2 - d1) / 3
    ana = dq[sel]
    scale = torch.maximum(ana.norm(dim=1), ana.norm(dim=1).median())
    rel = ((fd - ana).norm(dim=1) / scale).max().item()
    say(f"[wave] finite differences: max relative error {rel:.2e}")
    assert rel < 1e-6


def test_batch_and_graph_with_wave_rows(oracle_mod):
    rng = np.random.default_rng(41)
    keys, queries = [], []
    for s in (5, 6):
        kk, qq, _, _ = workloads.shell(seed=s, n_keys=1200, n_queries=250)
        keys.append(kk)
        queries.append(qq)
    for _ in range(3):
        side = rng.uniform(0.5, 2.0, 3)
        keys.append(rng.random((int(rng.integers(300, 2000)), 3)) * side)
        queries.append(rng.random((int(rng.integers(100, 900)), 3)) * side)
    b = natinterp3d_amd.BatchInterpolator(keys)
    b.set_timing(True)
    Ws = b.get_weights(queries)
    for p in range(len(keys)):
        single = natinterp3d_amd.Interpolator(keys[p])
        Wo = oracle_mod.OracleMesh(single.mesh_and_vertices.export()).get_weights(
            np.ascontiguousarray(queries[p], np.float64) - single._centroid)
        assert csr_equal(Ws[p], Wo), f"problem {p}"
    # graph replays through the wave level equal eager
    dev = torch.device("cuda", 0)
    qc, off = b._centre(queries)
    qd = torch.from_numpy(qc).to(dev)
    crow0, col0, val0 = b.weights_device(qd, off)
    graph = natinterp3d_amd.WeightsGraph(b, qd, off)
    for _ in range(3):
        crow, col, val = graph.run()
        torch.cuda.synchronize()
        assert torch.equal(crow, crow0) and torch.equal(col, col0) and torch.equal(val, val0)
