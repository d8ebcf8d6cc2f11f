"""Backward pass of the differentiable interpolation (SibsonInterpolation):
dL/dvalues against W^T g from the forward CSR, dL/dqueries against central
finite differences of the oracle's CPU rows, W(q +- h e_k) @ values, at
h = 1e-6 and h/2, Richardson-combined ((4 D(h/2) - D(h)) / 3) so that the
h^2 truncation term cancels: next to dense clusters (key spacing ~1e-3) the
plain h = 1e-6 difference is itself off by ~1e-4 relative.

Tolerance: the query gradient must match to 1e-6 relative per query, the
scale being the larger of the row's own gradient and the median row's (the
differences carry an absolute rounding error of ~1e-14 / h from the weights,
which a row with a small gradient cannot hide); W^T g to 1e-12 relative
(only the order of the float64 additions differs).
"""
import numpy as np
import pytest

import natinterp3d_amd
from natinterp3d_amd import workloads

pytestmark = pytest.mark.gpu

H = 1e-6
FD_RTOL = 1e-6


def say(msg):
    print(msg, flush=True)


def central_diff(om, qc, values, g, h):
    """Central differences of sum_d g[i,d] (W(q_i) @ values)[d] w.r.t. q_i,
    from the oracle's rows."""
    out = np.zeros((len(qc), 3))
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        yp = om.get_weights(qc + e, num_threads=8) @ values
        ym = om.get_weights(qc - e, num_threads=8) @ values
        out[:, k] = ((yp - ym) * g).sum(1) / (2 * h)
    return out


def fd_grad(om, qc, values, g, h=H):
    d1 = central_diff(om, qc, values, g, h)
    d2 = central_diff(om, qc, values, g, h / 2)
    return (4 * d2 - d1) / 3


def run_backward(interp, queries, values, g):
    import torch
    dev = torch.device("cuda", 0)
    qt = torch.tensor(queries, device=dev, requires_grad=True)
    vt = torch.tensor(values, device=dev, requires_grad=True)
    y = interp.interpolate_torch(qt, vt)
    (y * torch.tensor(g, device=dev, dtype=y.dtype)).sum().backward()
    return (y.detach().cpu().numpy(), qt.grad.cpu().numpy(), vt.grad.cpu().numpy())


def check_fd(dq, fd, what):
    scale = np.abs(fd).max(1)
    rel = np.abs(dq - fd).max(1) / np.maximum(scale, np.median(scale))
    worst = int(np.argmax(rel))
    say(f"[{what}] worst relative query-gradient error {rel[worst]:.2e} (row {worst}, "
        f"|fd| {scale[worst]:.3g}); median {np.median(rel):.2e}")
    assert rel[worst] <= FD_RTOL, f"{what}: row {worst} dq {dq[worst]} fd {fd[worst]}"


def problem(oracle_mod, keys, queries, values, label):
    keys = np.ascontiguousarray(keys, np.float64)
    say(f"[{label}] N={len(keys)} M={len(queries)} build")
    interp = natinterp3d_amd.Interpolator(keys)
    om = oracle_mod.OracleMesh(interp.mesh_and_vertices.export())
    c = interp._centroid
    qc = np.ascontiguousarray(queries, np.float64) - c
    return interp, om, qc


def test_uniform_values_and_queries(oracle_mod):
    rng = np.random.default_rng(11)
    keys = rng.random((3000, 3))
    queries = 0.05 + 0.9 * rng.random((2000, 3))
    values = rng.standard_normal((3000, 3))
    g = rng.standard_normal((2000, 3))
    interp, om, qc = problem(oracle_mod, keys, queries, values, "uniform")
    say("[uniform] forward + backward")
    y, dq, dv = run_backward(interp, queries, values, g)
    W = om.get_weights(qc)
    # forward: the same product as the numpy path (float64 W, float64 values)
    assert np.array_equal(y, interp.interpolate(queries, values))
    ref_dv = W.T @ g
    np.testing.assert_allclose(dv, ref_dv, rtol=1e-12, atol=1e-12 * np.abs(ref_dv).max())
    say("[uniform] finite differences")
    check_fd(dq, fd_grad(om, qc, values, g), "uniform")


def test_scalar_values_float32(oracle_mod):
    """values [N] float32 and float32 queries: results and gradients in
    float32, computed from float64 weights."""
    import torch
    rng = np.random.default_rng(12)
    keys = rng.random((1500, 3))
    queries = (0.1 + 0.8 * rng.random((500, 3))).astype(np.float32)
    values = rng.standard_normal(1500).astype(np.float32)
    interp = natinterp3d_amd.Interpolator(keys)
    dev = torch.device("cuda", 0)
    qt = torch.tensor(queries, device=dev, requires_grad=True)
    vt = torch.tensor(values, device=dev, requires_grad=True)
    y = interp.interpolate_torch(qt, vt)
    assert y.dtype == torch.float32 and y.shape == (500,)
    y.sum().backward()
    assert qt.grad.dtype == torch.float32 and vt.grad.dtype == torch.float32
    om = oracle_mod.OracleMesh(interp.mesh_and_vertices.export())
    qc = queries.astype(np.float64) - interp._centroid
    W = om.get_weights(qc)
    np.testing.assert_allclose(y.detach().cpu().numpy(), W @ values.astype(np.float64),
                               rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(vt.grad.cpu().numpy(), np.asarray(W.sum(0)).ravel(),
                               rtol=1e-5, atol=1e-6)
    fd = fd_grad(om, qc, values.astype(np.float64)[:, None], np.ones((500, 1)))
    np.testing.assert_allclose(qt.grad.cpu().numpy(), fd, rtol=1e-4,
                               atol=1e-4 * np.abs(fd).max())


def test_shell_queries(oracle_mod):
    """Queries inside a hollow noisy shell: hundreds of natural neighbours per
    row, every row made by the larger-capacity levels (no topology record)."""
    keys, queries, values, _ = workloads.shell(n_keys=20_000, n_queries=4_000)
    rng = np.random.default_rng(13)
    g = rng.standard_normal((len(queries), 3))
    interp, om, qc = problem(oracle_mod, keys, queries, values, "shell")
    say("[shell] forward + backward")
    y, dq, dv = run_backward(interp, queries, values, g)
    W = om.get_weights(qc, num_threads=8)
    np.testing.assert_allclose(y, W @ values, rtol=0, atol=0)
    ref_dv = W.T @ g
    np.testing.assert_allclose(dv, ref_dv, rtol=1e-12, atol=1e-12 * np.abs(ref_dv).max())
    sel = np.random.default_rng(14).choice(len(queries), 1000, replace=False)
    say("[shell] finite differences on 1000 queries")
    check_fd(dq[sel], fd_grad(om, qc[sel], values, g[sel]), "shell")


def test_mixed_levels_and_special_rows(oracle_mod):
    """Blob keys, queries outside the data's hull (empty rows, as in the
    reference), queries on data points (coincident rows) and far-away
    queries (outside the super simplex: empty rows)."""
    rng = np.random.default_rng(15)
    centers = rng.random((4, 3))
    keys = np.vstack([rng.random((2500, 3)),
                      centers[rng.integers(0, 4, 1500)] + 0.004 * rng.standard_normal((1500, 3))])
    keys = keys[rng.permutation(len(keys))]
    queries = np.vstack([0.02 + 0.96 * rng.random((1500, 3)),
                         centers[rng.integers(0, 4, 300)] + 0.004 * rng.standard_normal((300, 3)),
                         1.1 + 0.2 * rng.random((50, 3)),  # outside the hull
                         keys[:3],                           # coincident
                         np.full((2, 3), 1e6)])              # outside the super simplex
    values = rng.standard_normal((len(keys), 2))
    g = rng.standard_normal((len(queries), 2))
    interp, om, qc = problem(oracle_mod, keys, queries, values, "mixed")
    y, dq, dv = run_backward(interp, queries, values, g)
    W = om.get_weights(qc)
    assert np.array_equal(y, W @ values)
    ref_dv = W.T @ g
    np.testing.assert_allclose(dv, ref_dv, rtol=1e-12, atol=1e-12 * np.abs(ref_dv).max())
    assert np.isfinite(dq).all()
    n_reg = 1800
    nnz = np.diff(W.indptr)
    assert (nnz[n_reg:n_reg + 50] == 0).all() and (nnz[n_reg + 53:] == 0).all()
    assert (nnz[n_reg + 50:n_reg + 53] == 1).all()
    assert (dq[n_reg:] == 0).all()  # empty and coincident rows
    check_fd(dq[:n_reg], fd_grad(om, qc[:n_reg], values, g[:n_reg]), "mixed")


def test_saved_topology_survives_other_calls(oracle_mod):
    """The backward of one forward is unaffected by another forward on the same
    mesh in between (each forward keeps its own saved topology)."""
    import torch
    rng = np.random.default_rng(16)
    keys = rng.random((2000, 3))
    q1 = 0.1 + 0.8 * rng.random((700, 3))
    q2 = 0.1 + 0.8 * rng.random((1300, 3))
    values = rng.standard_normal((2000, 1))
    interp = natinterp3d_amd.Interpolator(keys)
    dev = torch.device("cuda", 0)
    vt = torch.tensor(values, device=dev)
    qt1 = torch.tensor(q1, device=dev, requires_grad=True)
    y1 = interp.interpolate_torch(qt1, vt)
    qt2 = torch.tensor(q2, device=dev, requires_grad=True)
    y2 = interp.interpolate_torch(qt2, vt)
    y1.sum().backward()
    y2.sum().backward()
    qt1b = torch.tensor(q1, device=dev, requires_grad=True)
    interp.interpolate_torch(qt1b, vt).sum().backward()
    assert torch.equal(qt1.grad, qt1b.grad)
    om = oracle_mod.OracleMesh(interp.mesh_and_vertices.export())
    qc2 = q2 - interp._centroid
    check_fd(qt2.grad.cpu().numpy(), fd_grad(om, qc2, values, np.ones((1300, 1))), "second")


def test_grad_queries_rejects_mismatched_topology():
    import ctypes

    import torch

    from natinterp3d_amd import _lib
    rng = np.random.default_rng(17)
    keys = rng.random((500, 3))
    interp = natinterp3d_amd.Interpolator(keys)
    m = interp.mesh_and_vertices
    dev = torch.device("cuda", 0)
    q = torch.tensor(0.2 + 0.6 * rng.random((100, 3)) - interp._centroid, device=dev)
    m.weights_device(q)
    L = _lib.lib()
    nb = ctypes.c_int64()
    assert L.nat_topology_save(m.handle, None, 0, ctypes.byref(nb), None) == 0
    topo = torch.empty(int(nb.value), dtype=torch.uint8, device=dev)
    assert L.nat_topology_save(m.handle, ctypes.c_void_p(topo.data_ptr()), 16,
                               ctypes.byref(nb), None) == -7
    assert L.nat_topology_save(m.handle, ctypes.c_void_p(topo.data_ptr()), topo.numel(),
                               ctypes.byref(nb), None) == 0
    gy = torch.zeros((99, 1), dtype=torch.float64, device=dev)
    vals = torch.zeros((500, 1), dtype=torch.float64, device=dev)
    dq = torch.zeros((99, 3), dtype=torch.float64, device=dev)
    rowptr = torch.zeros(100, dtype=torch.int32, device=dev)
    rc = L.nat_grad_queries(m.handle, ctypes.c_void_p(topo.data_ptr()),
                            ctypes.c_void_p(q.data_ptr()), 99, ctypes.c_void_p(rowptr.data_ptr()),
                            None, ctypes.c_void_p(gy.data_ptr()), ctypes.c_void_p(vals.data_ptr()),
                            1, ctypes.c_void_p(dq.data_ptr()), None)
    assert rc == -1 and "100 queries" in _lib.last_error()


def test_topology_target_in_place_and_fallbacks():
    """Records written straight into the saved topology (nat_topology_target)
    give the same values, products and gradients as a saved copy; a target too
    small for the records, or set for another number of queries, falls back to
    the copy."""
    import ctypes
    import torch
    from natinterp3d_amd import _lib, autograd
    k, q, v = workloads.uniform_small(seed=9, n_keys=3000, n_queries=5000)
    interp = natinterp3d_amd.Interpolator(k)
    mesh = interp.mesh_and_vertices
    key = mesh.handle.value
    dev = torch.device("cuda", 0)
    vals = torch.tensor(np.stack([v, -2 * v], 1), device=dev)
    gy = torch.randn(len(q), 2, dtype=torch.float64, device=dev,
                     generator=torch.Generator(device=dev).manual_seed(3))

    def run(bytes_per_query):
        if bytes_per_query is None:
            autograd._TOPO_BYTES_PER_QUERY.pop(key, None)
        else:
            autograd._TOPO_BYTES_PER_QUERY[key] = bytes_per_query
        qt = torch.tensor(q, device=dev, requires_grad=True)
        vt = vals.clone().requires_grad_(True)
        y = interp.interpolate_torch(qt, vt)
        dq, dv = torch.autograd.grad(y, (qt, vt), gy)
        return y, dq, dv

    def same(got, ref, what):
        # y and dL/dq bit for bit; dL/dvalues to 1e-12 (float64 atomics: the
        # order of the additions into one output is not fixed)
        assert torch.equal(got[0], ref[0]), what
        assert torch.equal(got[1], ref[1]), what
        torch.testing.assert_close(got[2], ref[2], rtol=1e-12, atol=1e-12)

    base = run(1e-3)  # target far too small: the records are copied
    for per_q in (None, 5000.0):  # default estimate; ample room: written in place
        same(run(per_q), base, per_q)
    # a target set for another query count is consumed without being written
    L = _lib.lib()
    buf = torch.full((1 << 22,), 7, dtype=torch.uint8, device=dev)
    assert L.nat_topology_target(mesh.handle, ctypes.c_void_p(buf.data_ptr()), buf.numel(),
                                 len(q) + 1) == 0
    crow, col, val = mesh.weights_device(torch.tensor(q - interp._centroid, device=dev))
    torch.cuda.synchronize()
    assert bool((buf == 7).all())
    assert torch.equal(torch.tensor(interp.get_weights(q).data, device=dev), val)
    same(run(1e-3), base, "after another-M target")
    assert L.nat_topology_target(mesh.handle, ctypes.c_void_p(buf.data_ptr() + 8), 64, 5) < 0
    say("[grad] topology target: in place, too small and other-M targets agree")


def test_shell_gradient_from_big_records_equals_recompute():
    """Rows of the wave-per-query level keep a big record when the forward
    saves its topology (autograd), and their gradient comes from it; a
    topology saved without them (a plain forward, then nat_topology_save)
    makes the backward search those cavities again.  Both give the same
    dL/dq (to rounding: the record path divides by the forward's row total,
    the other by its own), and the record path's topology is the larger."""
    import ctypes

    import torch

    from natinterp3d_amd import _lib
    keys, queries, values, _ = workloads.shell(n_keys=20_000, n_queries=3_000)
    interp = natinterp3d_amd.Interpolator(np.ascontiguousarray(keys, np.float64))
    mesh = interp.mesh_and_vertices
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(19)
    g = rng.standard_normal((len(queries), 3))
    # the record path (autograd)
    qt = torch.tensor(queries, device=dev, requires_grad=True)
    vt = torch.tensor(values, device=dev)
    y = interp.interpolate_torch(qt, vt)
    (y * torch.tensor(g, device=dev)).sum().backward()
    dq_rec = qt.grad.cpu().numpy()
    assert mesh.timings()["wave_items"] >= 0.9 * len(queries)  # rows of the wave level
    # the search-again path: a plain forward, its topology saved afterwards
    L = _lib.lib()
    qc = torch.tensor(np.ascontiguousarray(queries - interp._centroid), device=dev)
    crow, col, val = mesh.weights_device(qc)
    nb = ctypes.c_int64()
    assert L.nat_topology_save(mesh.handle, None, 0, ctypes.byref(nb), None) == 0
    topo = torch.empty(int(nb.value), dtype=torch.uint8, device=dev)
    assert L.nat_topology_save(mesh.handle, ctypes.c_void_p(topo.data_ptr()), topo.numel(),
                               ctypes.byref(nb), None) == 0
    vals = torch.tensor(values, device=dev, dtype=torch.float64)
    gy = torch.tensor(g, device=dev, dtype=torch.float64)
    dq = torch.zeros((len(queries), 3), dtype=torch.float64, device=dev)
    rc = L.nat_grad_queries(mesh.handle, ctypes.c_void_p(topo.data_ptr()),
                            ctypes.c_void_p(qc.data_ptr()), len(queries),
                            ctypes.c_void_p(crow.data_ptr()), ctypes.c_void_p(val.data_ptr()),
                            ctypes.c_void_p(gy.data_ptr()), ctypes.c_void_p(vals.data_ptr()), 3,
                            ctypes.c_void_p(dq.data_ptr()), None)
    assert rc == 0, _lib.last_error()
    torch.cuda.synchronize()
    dq_again = dq.cpu().numpy()
    scale = np.maximum(np.abs(dq_again).max(1), np.median(np.abs(dq_again).max(1)))
    rel = np.abs(dq_rec - dq_again).max(1) / scale
    say(f"[big records] max relative difference {rel.max():.2e}; saved topology "
        f"{mesh.timings()['wave_items']} wave rows, {int(nb.value):,} bytes without records")
    assert rel.max() <= 1e-9
    # the autograd topology held the big records: far more bytes per query
    from natinterp3d_amd import autograd
    per_query_rec = autograd._TOPO_BYTES_PER_QUERY[mesh.handle.value]
    assert per_query_rec > 4 * int(nb.value) / len(queries)
