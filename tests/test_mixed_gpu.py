"""Mixed-precision rows: neighbour sets (CSR structure) identical to the FP64
path; weights within 1e-5 of the FP64 weights (the bound the mixed path
promises; measured errors are printed); the path counters add up."""
import numpy as np
import pytest

from natinterp3d_amd import workloads
from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP

pytestmark = pytest.mark.gpu

TOL = 1e-5


def say(msg):
    print(msg, flush=True)


def compare(keys, queries, label, oracle_mod=None):
    keys = np.ascontiguousarray(keys, np.float64)
    c = keys.mean(axis=0)
    kc = np.ascontiguousarray(keys - c)
    qc = np.ascontiguousarray(queries, np.float64) - c
    m = MeshAndVerticesHIP(kc, device=0)
    W64 = m.get_weights(qc)
    m.set_precision("mixed")
    Wmx = m.get_weights(qc)
    t = m.timings()
    assert np.array_equal(W64.indptr, Wmx.indptr), f"{label}: row lengths differ"
    assert np.array_equal(W64.indices, Wmx.indices), f"{label}: neighbour order differs"
    err = np.abs(W64.data - Wmx.data).max() if W64.nnz else 0.0
    n32, n64 = t["mixed_fp32"], t["mixed_fp64"]
    say(f"[{label}] max |w_mixed - w_fp64| = {err:.3g}; tets+faces in FP32 {n32:,}, "
        f"recomputed in FP64 {n64:,} ({100.0 * n64 / max(1, n32 + n64):.2f}%)")
    assert err <= TOL, f"{label}: mixed weights off by {err}"
    assert n32 > 0
    if oracle_mod is not None:
        Wo = oracle_mod.OracleMesh(m.export()).get_weights(qc)
        assert np.abs(Wo.data - Wmx.data).max() <= TOL
    m.set_precision("fp64")
    W2 = m.get_weights(qc)
    assert np.array_equal(W2.data, W64.data), f"{label}: FP64 not restored"
    return err, n32, n64


def test_mixed_uniform(oracle_mod):
    k, q, _ = workloads.uniform_small(n_keys=5000, n_queries=20000)
    compare(k, q, "uniform", oracle_mod)


def test_mixed_blobs():
    k, q, _ = workloads.plain_at_scale(n_keys=50_000, n_queries=50_000)
    compare(k, q, "blobs")


def test_mixed_lattice_near_cospherical():
    k, q, _ = workloads.lattice_jitter(n_side=25, n_queries=30000)
    err, n32, n64 = compare(k, q, "lattice")
    assert n64 > 0  # near-degenerate tets are recomputed in FP64


def test_predicate_audit():
    import torch
    # well-separated random data: every decision certified
    k, q, _ = workloads.uniform_small(n_keys=3000, n_queries=5000)
    c = k.mean(axis=0)
    m = MeshAndVerticesHIP(np.ascontiguousarray(k - c), device=0)
    dec, unc = m.predicate_audit(torch.tensor(q - c, device="cuda"))
    say(f"[audit uniform] {dec:,} decisions, {unc} uncertain")
    assert dec > 5000 * 40 and unc == 0
    # queries on the circumsphere of a tet, on the ray from its circumcentre
    # through its centroid (away from its vertices, where the determinant's
    # terms would all shrink with it): that tet's test is decided by rounding
    e = m.export()
    V, tv = e["verts"], e["tet_v"]
    qs = []
    for t in range(400):
        if (tv[t] >= 3000).any():
            continue
        P = V[tv[t]]
        A = 2.0 * (P[1:] - P[0])
        b = (P[1:] ** 2).sum(1) - (P[0] ** 2).sum()
        cc = np.linalg.solve(A, b)
        u = P.mean(axis=0) - cc
        qs.append(cc + np.linalg.norm(P[0] - cc) * u / np.linalg.norm(u))
    dec, unc = m.predicate_audit(torch.tensor(np.array(qs), device="cuda"))
    say(f"[audit on circumspheres] {dec:,} decisions, {unc} uncertain")
    assert unc > 0


def test_precision_argument_checks():
    rng = np.random.default_rng(0)
    m = MeshAndVerticesHIP(np.ascontiguousarray(rng.random((50, 3)) - 0.5), device=0)
    with pytest.raises(ValueError):
        m.set_precision("fp16")
    with pytest.raises(RuntimeError):
        m.set_precision("mixed", cond_threshold=0.0)
