"""Pinning the oracle (oracle/sibson_oracle.c) before trusting it.

1. Against the reference's own outputs: the golden fixtures written from the
   compiled reference (tests/golden/), and live against oracle/_ref when it
   is built here.
2. Against what the mathematics fixes: exact determinants on integer
   coordinates, circumcentre equidistance, the Voronoi-subcell partition of a
   tetrahedron's volume, Sibson-coordinate invariants, and natural-neighbour
   sets from an independent Delaunay triangulation (Qhull).
"""
import hashlib
import json
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.spatial

from tests.golden.cases import CASES, config1

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _mesh_from_golden(meshes, name):
    return {k: meshes[f"{name}__{k}"] for k in
            ("verts", "tet_v", "tet_n", "kd_left", "kd_right", "kd_pos", "kd_rect")}


def _csr_eq(W, indptr, indices, data):
    assert np.array_equal(W.indptr, indptr)
    assert np.array_equal(W.indices, indices)
    same = (W.data == data) | (np.isnan(W.data) & np.isnan(data))
    assert same.all()


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_reference_fixture(oracle_mod, name):
    """Oracle on the reference's mesh == the reference's CSR, bit for bit."""
    small = np.load(os.path.join(GOLDEN, "small_cases.npz"))
    meshes = np.load(os.path.join(GOLDEN, "meshes.npz"))
    # centre exactly as the API does: the mean of the C-contiguous copy
    # (natinterp3d.py:88-90; NumPy's pairwise sum depends on memory order)
    keys = np.ascontiguousarray(small[f"{name}__keys"])
    q = small[f"{name}__queries"] - keys.mean(axis=0)
    om = oracle_mod.OracleMesh(_mesh_from_golden(meshes, name))
    for threads in (0, 3):
        W = om.get_weights(q, num_threads=threads)
        _csr_eq(W, small[f"{name}__indptr"], small[f"{name}__indices"], small[f"{name}__data"])


def test_oracle_config1_digest(oracle_mod):
    """Configuration 1 (100,000 rows): SHA-256 of indptr/indices/data equal the
    reference's, and 50 sampled rows match exactly."""
    from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP
    gold = json.load(open(os.path.join(GOLDEN, "config1.json")))
    keys, q = config1()
    c = keys.mean(axis=0)
    m = MeshAndVerticesHIP(np.ascontiguousarray(keys - c), device=-1)
    W = oracle_mod.OracleMesh(m.export()).get_weights(q - c, num_threads=8)
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()  # noqa
    for k in ("indptr", "indices", "data"):
        assert sha(getattr(W, k)) == gold["sha256"][k], k
    for i, row in gold["rows"].items():
        i = int(i)
        sl = slice(W.indptr[i], W.indptr[i + 1])
        assert W.indices[sl].tolist() == row["indices"]
        assert [float(x).hex() for x in W.data[sl]] == row["data"]


@pytest.mark.parametrize("seed,n,m", [(0, 500, 2000), (1, 3000, 5000), (2, 64, 1000)])
def test_oracle_vs_compiled_reference(oracle_mod, ref_mod, seed, n, m):
    rng = np.random.default_rng(seed)
    keys = rng.standard_normal((n, 3))
    q = 1.5 * rng.standard_normal((m, 3))  # interior, hull-adjacent and exterior rows
    kc = np.ascontiguousarray(keys - keys.mean(0))
    qc = q - keys.mean(0)
    r = ref_mod.RefMesh(kc)
    om = oracle_mod.OracleMesh(r.export())
    Wr = r.get_weights(qc, num_threads=4)
    W = om.get_weights(qc, num_threads=4)
    _csr_eq(W, Wr.indptr, Wr.indices, Wr.data)


@pytest.mark.parametrize("label", ["plain_at_scale_100k", "lattice_jitter_50"])
def test_oracle_vs_compiled_reference_on_benchmark_recipes(oracle_mod, ref_mod, label):
    """The oracle's rows equal the compiled reference's on the two large
    benchmark recipes (the blob mix at 100k keys; the jittered 50^3 lattice,
    whose perturbed coordinates and near-co-spherical insphere tests are
    where a restatement would first drift), 200k queries each."""
    from natinterp3d_amd import workloads
    if label.startswith("plain"):
        keys, q, _ = workloads.plain_at_scale(n_keys=100_000, n_queries=200_000)
    else:
        keys, q, _ = workloads.lattice_jitter(n_side=50, n_queries=200_000)
    c = keys.mean(0)
    kc = np.ascontiguousarray(keys - c)
    qc = np.ascontiguousarray(q - c)
    r = ref_mod.RefMesh(kc)
    om = oracle_mod.OracleMesh(r.export())  # export before the reference packs its mesh
    Wr = r.get_weights(qc, num_threads=8)
    W = om.get_weights(qc, num_threads=8)
    _csr_eq(W, Wr.indptr, Wr.indices, Wr.data)


# ---------------------------------------------------------------- mathematics

def _det3(m):
    return (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1])
            - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
            + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))


def _det4(m):
    s = 0
    for j in range(4):
        minor = [[m[i][k] for k in range(4) if k != j] for i in range(1, 4)]
        s += (-1) ** j * m[0][j] * _det3(minor)
    return s


def test_orient3d_is_exact_determinant_on_small_integers(oracle_mod):
    rng = np.random.default_rng(3)
    for _ in range(500):
        a, b, c, d = rng.integers(-50, 51, (4, 3)).astype(float)
        exact = _det3([[int(a[i] - d[i]) for i in range(3)], [int(b[i] - d[i]) for i in range(3)],
                       [int(c[i] - d[i]) for i in range(3)]])
        assert oracle_mod.orient3d(a, b, c, d) == exact


def test_insphere_is_exact_lifted_determinant_on_small_integers(oracle_mod):
    rng = np.random.default_rng(4)
    for _ in range(500):
        p = rng.integers(-30, 31, (5, 3)).astype(float)
        e = p[4]
        rows = []
        for i in range(4):
            d = [int(p[i][k] - e[k]) for k in range(3)]
            rows.append(d + [d[0] ** 2 + d[1] ** 2 + d[2] ** 2])
        exact = _det4(rows)
        assert oracle_mod.insphere(*p[:4], e) == exact
        # and the orientation-swapped form used for negative tets flips it
        assert oracle_mod.insphere(p[1], p[0], p[2], p[3], e) == -exact
    # meaning of the sign: > 0 inside the circumsphere of a tet with
    # orient3d > 0, < 0 outside
    for _ in range(200):
        P = rng.standard_normal((4, 3))
        if oracle_mod.orient3d(*P) <= 0:
            P = P[[1, 0, 2, 3]]
        ct = oracle_mod.cc_tet(*P)
        r = np.linalg.norm(P[0] - ct)
        assert oracle_mod.insphere(*P, ct) > 0
        assert oracle_mod.insphere(*P, ct + np.array([3 * r, 0, 0])) < 0


def test_circumcentres_are_equidistant(oracle_mod):
    rng = np.random.default_rng(5)
    for _ in range(200):
        a, b, c, d = rng.standard_normal((4, 3))
        t = oracle_mod.cc_tri(a, b, c)
        r = [np.linalg.norm(t - x) for x in (a, b, c)]
        assert np.allclose(r, r[0], rtol=1e-9)
        assert abs(np.dot(t - a, np.cross(b - a, c - a))) < 1e-9  # in the plane
        s = oracle_mod.cc_tet(a, b, c, d)
        r = [np.linalg.norm(s - x) for x in (a, b, c, d)]
        assert np.allclose(r, r[0], rtol=1e-8)
    # degenerate fallbacks: the centroids
    a, b = np.zeros(3), np.array([1.0, 0, 0])
    assert np.array_equal(oracle_mod.cc_tri(a, b, 2 * b), (a + b + 2 * b) / 3.0)
    d = np.array([0.0, 1, 0])
    assert np.allclose(oracle_mod.cc_tet(a, b, d, b + d), (a + b + d + b + d) * 0.25)


def test_subcells_partition_the_tetrahedron(oracle_mod):
    """The four Voronoi subcells of a well-centred tetrahedron fill it."""
    rng = np.random.default_rng(6)
    done = 0
    while done < 100:
        P = rng.standard_normal((4, 3))
        ct = oracle_mod.cc_tet(*P)
        # well-centred: circumcentre strictly inside
        bary = np.linalg.solve(np.vstack([P.T, np.ones(4)]), np.append(ct, 1.0))
        if bary.min() < 0.05:
            continue
        fcc = [oracle_mod.cc_tri(*[P[j] for j in range(4) if j != k]) for k in range(4)]
        total = 0.0
        for k in range(4):
            o = sorted(j for j in range(4) if j != k)
            total += oracle_mod.subcell(P[k], P[o[0]], P[o[1]], P[o[2]], ct,
                                        fcc[o[2]], fcc[o[1]], fcc[o[0]])
        vol = abs(np.dot(P[1] - P[0], np.cross(P[2] - P[0], P[3] - P[0]))) / 6.0
        assert total == pytest.approx(vol, rel=1e-10)
        done += 1


def test_natural_neighbours_match_qhull(oracle_mod):
    """Interior queries: the row's column set equals the Delaunay neighbours
    of q in an independent triangulation of keys + {q}."""
    from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP
    rng = np.random.default_rng(7)
    keys = rng.random((400, 3))
    qs = 0.3 + 0.4 * rng.random((100, 3))
    c = keys.mean(0)
    m = MeshAndVerticesHIP(np.ascontiguousarray(keys - c), device=-1)
    W = oracle_mod.OracleMesh(m.export()).get_weights(qs - c)
    for i, q in enumerate(qs):
        tri = scipy.spatial.Delaunay(np.vstack([keys, q]))
        ptr, nb = tri.vertex_neighbor_vertices
        expect = set(nb[ptr[400]:ptr[401]].tolist())
        got = set(W.indices[W.indptr[i]:W.indptr[i + 1]].tolist())
        assert got == expect, i


def test_sibson_invariants(oracle_mod):
    """Partition of unity, non-negativity and linear precision of the rows
    (tests/test_natinterp3d.py:18-114), on the oracle."""
    from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP
    rng = np.random.RandomState(12)
    keys = rng.randn(500, 3)
    q = keys.mean(0) + rng.randn(200, 3) * keys.std(0) * 0.24
    c = keys.mean(0)
    m = MeshAndVerticesHIP(np.ascontiguousarray(keys - c), device=-1)
    W = oracle_mod.OracleMesh(m.export()).get_weights(q - c)
    s = np.asarray(W.sum(axis=1)).ravel()
    inside = s > 0.5
    assert inside.sum() > 150
    np.testing.assert_allclose(s[inside], 1.0, atol=1e-10)
    assert W.min() >= -1e-12
    np.testing.assert_allclose((W @ keys)[inside], q[inside], atol=1e-8)


def test_oracle_plants_are_caught(oracle_mod):
    """A mistake planted in the checker's inputs must fail the fixture check:
    moving one data point by 1e-12 changes the rows."""
    small = np.load(os.path.join(GOLDEN, "small_cases.npz"))
    meshes = np.load(os.path.join(GOLDEN, "meshes.npz"))
    name = "linear"
    mesh = _mesh_from_golden(meshes, name)
    mesh["verts"] = mesh["verts"].copy()
    W0 = oracle_mod.OracleMesh(mesh).get_weights(
        small[f"{name}__queries"] - np.ascontiguousarray(small[f"{name}__keys"]).mean(0))
    mesh["verts"][W0.indices[0], 0] += 1e-12
    keys = np.ascontiguousarray(small[f"{name}__keys"])
    W = oracle_mod.OracleMesh(mesh).get_weights(small[f"{name}__queries"] - keys.mean(0))
    with pytest.raises(AssertionError):
        _csr_eq(W, small[f"{name}__indptr"], small[f"{name}__indices"], small[f"{name}__data"])


def test_exact_cavity_oracle_is_pinned(oracle_mod):
    """The checker of the robust predicate mode (oracle/exact_predicates.py):
    on a jittered lattice, (a) for random queries its neighbour sets equal an
    independent Delaunay triangulation's (Qhull on the mesh's vertices + q);
    (b) for queries on tet circumspheres its depth-first cavity equals the
    brute-force set of all tets holding q strictly inside (exact tests);
    (c) there it differs from the FP64 cavity on some queries -- the
    decisions the robust mode changes -- and agrees where the FP64 search's
    tests are all certified."""
    from natinterp3d_amd import workloads
    from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP
    from oracle import exact_predicates as X
    from tests.robust_cases import borderline_queries
    keys, _, _ = workloads.lattice_jitter(n_side=10, n_queries=1)
    kc = np.ascontiguousarray(keys - keys.mean(0))
    pk = MeshAndVerticesHIP(kc, device=-1).export()
    om = oracle_mod.OracleMesh(pk)
    n = pk["verts"].shape[0] - 4
    rng = np.random.default_rng(8)
    qr = (rng.random((60, 3)) - 0.5) * 0.6
    starts = om.locate(qr)
    for i, q in enumerate(qr):
        tri = scipy.spatial.Delaunay(np.vstack([pk["verts"][:n], q]))
        ptr, nb = tri.vertex_neighbor_vertices
        assert X.exact_neighbours(pk, q, starts[i]) == set(nb[ptr[n]:ptr[n + 1]].tolist()), i
    qb = borderline_queries(pk, 200, seed=3)
    starts = om.locate(qb)
    differ = 0
    for i, q in enumerate(qb):
        cav, _ = X.exact_cavity(pk, q, starts[i])
        if i < 60:
            assert sorted(cav) == X.in_sphere_brute(pk, q), i
        if set(cav) != set(om.cavity(q).tolist()):
            differ += 1
    assert differ > 0
