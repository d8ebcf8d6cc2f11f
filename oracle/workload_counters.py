"""Per-query work counters of a benchmark workload, from the oracle.

TEST/MEASUREMENT INFRASTRUCTURE: the oracle runs on a sample of the
workload's queries over the workload's mesh and reports the mean number of
k-d nodes entered, tets and vertices read, insphere tests, cavity tets,
boundary faces, subcell volumes and row entries per query.  These are the
inputs of the algorithmic byte and flop counts of the row kernel.

    python oracle/workload_counters.py [--sample 200000]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def plain_at_scale(n_keys=1_000_000, n_queries=10_000_000, seed=1):
    """1,000,000 keys (90% uniform in [0,1]^3, 10% in 20 isotropic Gaussian
    blobs of sigma 0.01 with uniform centres), permuted; 10,000,000 queries
    uniform in [0.02,0.98]^3; values float32 [N,3] ~ N(0,1)."""
    rng = np.random.default_rng(seed)
    n_u = n_keys * 9 // 10
    n_b = n_keys - n_u
    a = rng.random((n_u, 3))
    centers = rng.random((20, 3))
    b = centers[rng.integers(0, 20, n_b)] + 0.01 * rng.standard_normal((n_b, 3))
    keys = np.vstack([a, b])
    keys = keys[rng.permutation(len(keys))]
    queries = 0.02 + 0.96 * rng.random((n_queries, 3))
    values = rng.standard_normal((n_keys, 3)).astype(np.float32)
    return keys, queries, values


def counters(keys, queries, threads):
    from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP
    from oracle import oracle
    keys = np.ascontiguousarray(keys, np.float64)
    c = keys.mean(axis=0)
    m = MeshAndVerticesHIP(np.ascontiguousarray(keys - c), device=-1)
    om = oracle.OracleMesh(m.export())
    _, ctr = om.get_weights(queries - c, num_threads=threads, counters=True)
    q = ctr["queries"]
    per = {k: ctr[k] / q for k in ("kd_nodes", "walk_tests", "insphere", "cavity", "boundary",
                                   "face_cc", "sub3", "sub4", "nnz", "tets_read",
                                   "verts_read")}
    per["max_cavity"] = ctr["max_cavity"]
    per["max_nnz"] = ctr["max_nnz"]
    per["sample"] = q
    return per


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sample", type=int, default=200_000)
    ap.add_argument("--threads", type=int, default=len(os.sched_getaffinity(0)))
    a = ap.parse_args()
    t = time.time()
    keys, queries, _ = plain_at_scale()
    per = counters(keys, queries[:a.sample], a.threads)
    per["seconds"] = time.time() - t
    print(json.dumps(per, indent=1))


if __name__ == "__main__":
    main()
