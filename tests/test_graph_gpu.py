"""Captured launch sequences (WeightsGraph): every replay gives the eager
result bit for bit -- replays repeated with eager work and synchronisations
between them, queries changed in place between replays -- for a single mesh
and a batch; a replay that outgrows the captured capacities reports it."""
import numpy as np
import pytest

import natinterp3d_amd
from natinterp3d_amd import workloads
from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def eq(a, b):
    import torch
    return all(torch.equal(x, y) for x, y in zip(a, b))


def test_graph_replays_equal_eager_single():
    import torch
    k, q, _ = workloads.uniform_small(n_keys=4000, n_queries=30000)
    c = k.mean(axis=0)
    m = MeshAndVerticesHIP(np.ascontiguousarray(k - c), device=0)
    q1 = torch.tensor(q - c, device="cuda")
    rng = np.random.default_rng(5)
    q2 = torch.tensor(0.1 + 0.8 * rng.random(q.shape) - c, device="cuda")
    qbuf = q1.clone()
    g = natinterp3d_amd.WeightsGraph(m, qbuf)
    ref1 = m.weights_device(q1)
    ref2 = m.weights_device(q2)
    for i in range(3):
        assert eq(g.run(), ref1), f"replay {i} on the first queries"
        m.weights_device(q2)  # eager work between replays (no reallocation)
        torch.cuda.synchronize()
    qbuf.copy_(q2)
    for i in range(2):
        assert eq(g.run(), ref2), f"replay {i} on the second queries"
    say("[graph single] 5 replays equal to eager")


def test_graph_replays_equal_eager_batch():
    import torch
    probs = workloads.batch_problems(n_problems=12, keys_range=(200, 3000),
                                     queries_range=(100, 3000))
    b = natinterp3d_amd.BatchInterpolator([k for k, _, _ in probs])
    qs = [np.asarray(q, np.float64) - c for (_, q, _), c in zip(probs, b._centroids)]
    off = np.zeros(len(qs) + 1, np.int64)
    off[1:] = np.cumsum([len(x) for x in qs])
    qd = torch.tensor(np.concatenate(qs), device="cuda")
    ref = b.weights_device(qd, off)
    qbuf = qd.clone()
    g = natinterp3d_amd.WeightsGraph(b, qbuf, off)
    for i in range(3):
        assert eq(g.run(), ref), f"batch replay {i}"
        b.weights_device(qd, off)
        torch.cuda.synchronize()
    say("[graph batch] 3 replays equal to eager")


def test_graph_capacity_overflow_is_reported():
    import torch
    # inside a hollow shell every row has hundreds of entries, far beyond
    # what a graph made on empty rows holds
    k, q, _, _ = workloads.shell(n_keys=3000, n_queries=2000)
    c = k.mean(axis=0)
    m = MeshAndVerticesHIP(np.ascontiguousarray(k - c), device=0)
    far = torch.full((2000, 3), 1e6, dtype=torch.float64, device="cuda")  # empty rows
    g = natinterp3d_amd.WeightsGraph(m, far)
    crow, col, val = g.run()
    assert col.numel() == 0
    far.copy_(torch.tensor(q - c, device="cuda"))
    with pytest.raises(natinterp3d_amd.GraphCapacityError):
        g.run()
    crow2, col2, val2 = m.weights_device(far)  # the eager path still works
    assert col2.numel() > 2000 * 100


def test_graph_refuses_replay_after_precision_change():
    """A graph captured under one row arithmetic is refused once the mesh's
    precision changes (its kernels keep the capture's arithmetic); switching
    back makes it replay again, equal to eager."""
    import torch
    k, q, _ = workloads.uniform_small(n_keys=3000, n_queries=20000)
    c = k.mean(axis=0)
    m = MeshAndVerticesHIP(np.ascontiguousarray(k - c), device=0)
    qd = torch.tensor(q - c, device="cuda")
    m.set_precision("mixed")
    ref_mixed = m.weights_device(qd)
    g = natinterp3d_amd.WeightsGraph(m, qd.clone())
    assert eq(g.run(), ref_mixed)
    m.set_precision("fp64")
    with pytest.raises(natinterp3d_amd.GraphPrecisionError):
        g.run()
    m.set_precision("mixed", cond_threshold=0.05)
    with pytest.raises(natinterp3d_amd.GraphPrecisionError):
        g.run()
    m.set_precision("mixed")
    ref_mixed = m.weights_device(qd)
    torch.cuda.synchronize()
    assert eq(g.run(), ref_mixed), "replay after switching back"
    say("[graph precision] replay refused after fp64/threshold switch, equal after switch back")
