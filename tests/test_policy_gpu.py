"""Row-path policies and the autotuner: which kernel makes a row (the 16-lane
group kernel or the thread-per-query levels, from which level, in which work
order, on which grid) must never change the row.  Every policy's CSR equals
the default policy's bit for bit, and the default's equals the oracle."""
import numpy as np
import pytest
import torch

import natinterp3d_amd
from natinterp3d_amd import workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def mixed_problems(seed):
    """Uniform boxes plus two hollow shells (rows far beyond the group
    kernel's LDS tables)."""
    rng = np.random.default_rng(seed)
    keys, queries = [], []
    for _ in range(5):
        nk = int(rng.integers(200, 3000))
        side = rng.uniform(0.5, 2.0, 3)
        keys.append(rng.random((nk, 3)) * side)
        queries.append(rng.random((int(rng.integers(100, 2000)), 3)) * side)
    for s in (5, 6):
        k, q, _, _ = workloads.shell(seed=s, n_keys=1500, n_queries=300)
        keys.append(k)
        queries.append(q)
    return keys, queries


def csr_equal(a, b):
    return (np.array_equal(a.indptr, b.indptr) and np.array_equal(a.indices, b.indices)
            and np.array_equal(a.data, b.data))


POLICIES = [
    dict(first_level=2),
    dict(wave_rows=False),
    dict(wave_rows=False, first_level=2),
    dict(work_sort=False),
    dict(topo_grid=2, geom_grid=3),
    dict(first_level=2, work_sort=False, topo_grid=4, geom_grid=2),
    dict(geom_lanes=1),
    dict(geom_lanes=1, work_sort=False, geom_grid=2),
]


def test_batch_policies_give_identical_rows(oracle_mod):
    keys, queries = mixed_problems(31)
    P = len(keys)
    b = natinterp3d_amd.BatchInterpolator(keys)
    assert b.policy() == {"budgets": [96] * P, "first_level": 1, "work_sort": True,
                          "topo_grid": 1, "geom_grid": 1, "wave_rows": True,
                          "geom_lanes": 2}
    ref = b.get_weights(queries)
    for p in (0, P - 1):  # a box and a shell against the oracle
        single = natinterp3d_amd.Interpolator(keys[p])
        Wo = oracle_mod.OracleMesh(single.mesh_and_vertices.export()).get_weights(
            np.ascontiguousarray(queries[p], np.float64) - single._centroid)
        assert csr_equal(ref[p], Wo), f"default policy, problem {p}"
    budget_sets = [[0] * P, [96] * (P - 2) + [0, 0], [48] * P, [1] * P,
                   [0, 96, 64, 0, 32, 0, 96]]
    for budgets in budget_sets:
        b.set_policy(budgets=budgets)
        for extra in [{}] + POLICIES:
            b.set_policy(**extra)
            Ws = b.get_weights(queries)
            for p in range(P):
                assert csr_equal(Ws[p], ref[p]), f"budgets {budgets} {extra}: problem {p}"
            b.set_policy(first_level=1, work_sort=True, topo_grid=1, geom_grid=1,
                         geom_lanes=2)
        say(f"[policy] budgets {budgets}: all rows equal")


def test_single_mesh_routed_rows_and_gradients():
    k, q, v = workloads.uniform_small(seed=7, n_keys=3000, n_queries=4000)
    interp = natinterp3d_amd.Interpolator(k)
    ref = interp.get_weights(q)
    dev = torch.device("cuda", 0)
    qt = torch.tensor(q, device=dev, requires_grad=True)
    vt = torch.tensor(np.stack([v, 2 * v], 1), device=dev, requires_grad=True)
    g = torch.randn(len(q), 2, dtype=torch.float64, device=dev)
    y0 = interp.interpolate_torch(qt, vt)
    dq0, dv0 = torch.autograd.grad(y0, (qt, vt), g)
    mesh = interp.mesh_and_vertices
    for pol in ({"budgets": [0]}, {"budgets": [0], "first_level": 2}, {"budgets": [24]},
                {"geom_lanes": 1}):
        mesh.set_policy(**pol)
        assert csr_equal(interp.get_weights(q), ref), pol
        y = interp.interpolate_torch(qt, vt)
        assert torch.equal(y, y0), pol
        dq, dv = torch.autograd.grad(y, (qt, vt), g)
        torch.testing.assert_close(dq, dq0, rtol=1e-12, atol=1e-12)
        torch.testing.assert_close(dv, dv0, rtol=1e-12, atol=1e-12)
        mesh.set_policy(budgets=[96], first_level=1, geom_lanes=2)
        say(f"[policy] single mesh {pol}: rows, products and gradients equal")


def test_geometry_lanes_give_identical_rows_at_scale():
    """The FP64 geometry kernel with two lanes per row (accumulations split
    over the lanes in the reference's order) against one thread per row, on
    100k keys (10% in Gaussian blobs) and 1M queries: the device CSR bit for
    bit."""
    keys, queries, _ = workloads.plain_at_scale(n_keys=100_000, n_queries=1_000_000, seed=5)
    c = keys.mean(0)
    mesh = natinterp3d_amd.MeshAndVerticesHIP(np.ascontiguousarray(keys - c), device=0)
    qd = torch.from_numpy(np.ascontiguousarray(queries - c)).to(torch.device("cuda", 0))
    out = {}
    for lanes in (2, 1, 2):
        mesh.set_policy(geom_lanes=lanes)
        assert mesh.policy()["geom_lanes"] == lanes
        crow, col, val = mesh.weights_device(qd)
        torch.cuda.synchronize()
        cur = (crow.cpu(), col.cpu(), val.cpu())
        if lanes in out:
            assert all(torch.equal(a, b) for a, b in zip(cur, out[lanes])), "not repeatable"
        out[lanes] = cur
    assert all(torch.equal(a, b) for a, b in zip(out[1], out[2]))
    say(f"[policy] geometry lanes 1 and 2: {int(out[2][0][-1])} entries equal")


def test_autotune_batch_keeps_rows_and_reports():
    keys, queries = mixed_problems(32)
    b = natinterp3d_amd.BatchInterpolator(keys)
    ref = b.get_weights(queries)
    rep = b.autotune(queries, repeats=2)
    say(f"[policy] tuner report {rep}; chosen {b.policy()}")
    assert rep["candidates"] >= 4
    assert 0 < rep["tuned_ms"] <= rep["default_ms"]
    assert rep["left_group"] > 0  # the shells leave the group kernel
    pol = b.policy()
    assert rep["routed_problems"] == sum(1 for x in pol["budgets"] if x == 0)
    assert pol["first_level"] in (1, 2) and 1 <= pol["topo_grid"] <= 4
    Ws = b.get_weights(queries)
    for p in range(len(keys)):
        assert csr_equal(Ws[p], ref[p]), f"after tuning: problem {p}"


def test_autotune_single_mesh_and_graph_replay():
    k, q, _ = workloads.uniform_small(seed=8, n_keys=2000, n_queries=3000)
    interp = natinterp3d_amd.Interpolator(k)
    ref = interp.get_weights(q)
    rep = interp.autotune(q, repeats=2)
    say(f"[policy] single-mesh tuner report {rep}")
    assert rep["candidates"] >= 4 and rep["tuned_ms"] <= rep["default_ms"]
    assert csr_equal(interp.get_weights(q), ref)
    # a graph reads the budgets and first level at replay: captured under the
    # default, replayed under a smaller budget (fewer records, the same
    # entries, more thread-per-query rows on persistent grids: within the
    # capture's capacities), rows unchanged
    mesh = interp.mesh_and_vertices
    dev = torch.device("cuda", 0)
    qd = torch.tensor(q - interp._centroid, device=dev)
    mesh.set_policy(budgets=[96], first_level=1, work_sort=True)
    graph = natinterp3d_amd.WeightsGraph(mesh, qd)
    mesh.set_policy(budgets=[40], first_level=2)
    for _ in range(2):
        crow, col, val = graph.run()
        torch.cuda.synchronize()
        assert np.array_equal(crow.cpu().numpy(), ref.indptr)
        assert np.array_equal(col.cpu().numpy(), ref.indices)
        assert np.array_equal(val.cpu().numpy(), ref.data)
    mesh.set_policy(budgets=[96], first_level=1)


def test_routed_rows_outside_the_hull_have_zero_gradient():
    """With every row routed past the group kernel (budget 0), queries outside
    the data's hull and outside the super simplex (empty rows) get a zero
    query gradient from the larger levels, and the other rows the same
    gradient as under the default policy."""
    rng = np.random.default_rng(41)
    keys = rng.random((1500, 3))
    q = np.vstack([0.05 + 0.9 * rng.random((400, 3)),
                   1.2 + 0.1 * rng.random((20, 3)),   # outside the hull
                   np.full((2, 3), 1e6)])               # outside the super simplex
    interp = natinterp3d_amd.Interpolator(keys)
    dev = torch.device("cuda", 0)
    vt = torch.tensor(rng.standard_normal((1500, 2)), device=dev)
    g = torch.tensor(rng.standard_normal((len(q), 2)), device=dev)

    def grads():
        qt = torch.tensor(q, device=dev, requires_grad=True)
        y = interp.interpolate_torch(qt, vt)
        (dq,) = torch.autograd.grad(y, (qt,), g)
        return y, dq

    y0, dq0 = grads()
    mesh = interp.mesh_and_vertices
    mesh.set_policy(budgets=[0])
    try:
        y1, dq1 = grads()
    finally:
        mesh.set_policy(budgets=[96])
    assert torch.equal(y0, y1)
    nnz = np.diff(interp.get_weights(q).indptr)
    assert (nnz[400:] == 0).all()
    assert bool((dq1[400:] == 0).all()) and bool(torch.isfinite(dq1).all())
    torch.testing.assert_close(dq1[:400], dq0[:400], rtol=1e-12, atol=1e-12)
