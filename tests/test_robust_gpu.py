"""The opt-in robust predicate mode on the GPU, on a jittered lattice.

Queries are of two kinds: random ones, and ones placed on the circumspheres
of mesh tets (tests/robust_cases.py), where the FP64 insphere sign is not
certified and often wrong.

  * default predicates: every row bit-identical to the FP64 oracle's (the
    reference's non-robust decisions), no exact evaluation counted;
  * robust predicates: each row's neighbour set equals the one the exact
    cavity search gives (oracle/exact_predicates.py, itself pinned against
    brute force and Qhull in tests/test_oracle.py), and, for random queries,
    the one of an independent Delaunay triangulation (Qhull on the mesh's
    vertices + q); rows that differ from the default ones are no more than
    the decisions the kernels report as changed; robust weights keep the
    partition of unity and linear precision and stay within 1e-9 of the
    default weights (a tet whose circumsphere passes through q adds or
    removes a neighbour of weight ~0);
  * every row path (16-lane group search, wave per query, thread per query)
    gives the same robust rows.

Queries on tet circumspheres make Qhull's neighbour sets ambiguous (it
resolves near-cospherical points with its own tolerances: 61 of 100 such
sets differed from the exact ones on CPU), so they are checked against the
exact oracle only.
"""
import numpy as np
import pytest
import scipy.spatial

from natinterp3d_amd import workloads
from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP
from tests.robust_cases import borderline_queries

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def row(W, i):
    a, b = W.indptr[i], W.indptr[i + 1]
    return W.indices[a:b], W.data[a:b]


def same_csr(A, B):
    return (np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
            and np.array_equal(A.data, B.data))


@pytest.fixture(scope="module")
def case(oracle_mod):
    keys, _, _ = workloads.lattice_jitter(n_side=16, n_queries=1)
    kc = np.ascontiguousarray(keys - keys.mean(0))
    m = MeshAndVerticesHIP(kc, device=0)
    pk = m.export()
    qb = borderline_queries(pk, 1500, seed=5)
    rng = np.random.default_rng(6)
    qr = (rng.random((20000, 3)) - 0.5) * 0.9
    Q = np.ascontiguousarray(np.vstack([qb, qr]))
    om = oracle_mod.OracleMesh(pk)
    return m, pk, om, Q, len(qb)


def test_default_predicates_stay_bit_exact(case):
    m, pk, om, Q, nb = case
    say(f"[robust] default predicates, {len(Q)} queries ({nb} on circumspheres)")
    W = m.get_weights(Q)
    assert same_csr(W, om.get_weights(Q))
    t = m.timings()
    assert t["robust_exact"] == 0 and t["robust_changed"] == 0


def test_robust_rows_equal_exact_neighbour_sets(case):
    from oracle import exact_predicates as X
    m, pk, om, Q, nb = case
    W = m.get_weights(Q)
    m.set_predicates("robust")
    try:
        Wr = m.get_weights(Q)
        t = m.timings()
    finally:
        m.set_predicates("fp64")
    say(f"[robust] exact evaluations {t['robust_exact']}, changed {t['robust_changed']}, "
        f"unresolved {t['robust_unresolved']}")
    assert t["robust_exact"] > 0 and t["robust_changed"] > 0 and t["robust_unresolved"] == 0
    differ = [i for i in range(len(Q))
              if not (np.array_equal(row(W, i)[0], row(Wr, i)[0])
                      and np.array_equal(row(W, i)[1], row(Wr, i)[1]))]
    set_differ = [i for i in differ if set(row(W, i)[0]) != set(row(Wr, i)[0])]
    say(f"[robust] rows changed {len(differ)}, neighbour sets changed {len(set_differ)}")
    assert 0 < len(set_differ) <= len(differ) <= t["robust_changed"]
    starts = om.locate(Q)
    # every changed row and every circumsphere query against the exact cavity
    check = sorted(set(differ) | set(range(nb)))
    for i in check:
        got = set(row(Wr, i)[0].tolist())
        assert got == X.exact_neighbours(pk, Q[i], starts[i]), f"query {i}"
    # random queries: exact oracle on a sample, Qhull on a smaller one
    n = pk["verts"].shape[0] - 4
    rng = np.random.default_rng(1)
    for i in rng.choice(np.arange(nb, len(Q)), 300, replace=False):
        assert set(row(Wr, i)[0].tolist()) == X.exact_neighbours(pk, Q[i], starts[i])
    for i in rng.choice(np.arange(nb, len(Q)), 40, replace=False):
        tri = scipy.spatial.Delaunay(np.vstack([pk["verts"][:n], Q[i]]))
        ptr, nbv = tri.vertex_neighbor_vertices
        assert set(row(Wr, i)[0].tolist()) == set(nbv[ptr[n]:ptr[n + 1]].tolist()), i
    # weights: partition of unity, linear precision, continuity with FP64
    V = pk["verts"][:n]
    for i in check:
        c, w = row(Wr, i)
        assert abs(w.sum() - 1.0) < 1e-12
        assert np.abs(w @ V[c] - Q[i]).max() < 1e-9
        d = dict(zip(row(W, i)[0].tolist(), row(W, i)[1].tolist()))
        r = dict(zip(c.tolist(), w.tolist()))
        gap = max(abs(d.get(k, 0.0) - r.get(k, 0.0)) for k in set(d) | set(r))
        assert gap < 1e-9, (i, gap)


@pytest.mark.parametrize("policy", [{"budgets": [0], "wave_rows": True},
                                    {"budgets": [0], "wave_rows": False}])
def test_robust_rows_agree_across_row_paths(case, policy):
    m, pk, om, Q, nb = case
    m.set_predicates("robust")
    try:
        Wq = m.get_weights(Q)
        m.set_policy(**policy)
        Wp = m.get_weights(Q)
    finally:
        m.set_policy(budgets=[96], wave_rows=True)
        m.set_predicates("fp64")
    say(f"[robust] policy {policy}: rows equal to the group kernel's")
    assert same_csr(Wq, Wp)


def test_mixed_error_statement(case):
    """The per-call error statement of mixed rows: with every row sampled it
    is exactly the largest |w_mixed - w_fp64| of the call, and sampling does
    not change the rows."""
    m, pk, om, Q, nb = case
    W = m.get_weights(Q)
    m.set_precision("mixed")
    try:
        m.set_mixed_check(0)
        Wm0 = m.get_weights(Q)
        assert m.timings()["mixed_err_rows"] == 0
        m.set_mixed_check(1)
        Wm1 = m.get_weights(Q)
        t1 = m.timings()
        m.set_mixed_check(1024)
        Wm = m.get_weights(Q)
        t = m.timings()
    finally:
        m.set_precision("fp64")
        m.set_mixed_check(1024)
    assert same_csr(Wm0, Wm1) and same_csr(Wm0, Wm)
    assert np.array_equal(Wm.indices, W.indices)
    err = float(np.abs(Wm.data - W.data).max())
    say(f"[mixed] max |dw| {err:.3g}; statement over all rows {t1['mixed_err_max']:.3g} "
        f"({t1['mixed_err_rows']} rows, {t1['mixed_err_entries']} entries); every 1024th: "
        f"{t['mixed_err_max']:.3g} ({t['mixed_err_rows']} rows)")
    assert t1["mixed_err_max"] == err
    assert 0.9 * len(Q) <= t1["mixed_err_rows"] <= len(Q)
    assert t1["mixed_err_entries"] <= W.nnz
    assert 0 < t["mixed_err_rows"] <= (len(Q) + 1023) // 1024
    assert t["mixed_err_max"] <= err
