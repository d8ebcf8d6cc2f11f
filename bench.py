#!/usr/bin/env python3
"""Sibson query throughput on one MI355X.

Workload (the plain-at-scale configuration): 1,000,000 keys (90% uniform in
[0,1]^3, 10% in 20 isotropic Gaussian blobs of sigma 0.01), 10,000,000
queries uniform in [0.02,0.98]^3, values float32 [N,3].  The mesh is built
once on the host (untimed, reported on stderr).  One step is the forward pass
over all queries with inputs resident in HBM: every weight row computed on
the GPU, the CSR assembled in device memory, and W @ values.

Prints one JSON line on stdout.  ``roofline`` is the row kernel's achieved
algorithmic bandwidth (bytes per query from the oracle's work counters of this
workload, times queries per launch, over the launch time measured with HIP
events on the launch stream) against MI355X's 8 TB/s.  ``cpu_baseline`` is the
compiled reference (oracle/_ref) on the host's cores over a bounded sample of
the same queries.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_OPS_PEAK = 78.6e12 / 2  # non-fused FP64 operations per second (FMA counts 2 flops)

# Mean work per query of the plain-at-scale workload, measured by the oracle's
# counters on its first 200,000 queries (oracle/workload_counters.py).
WORK_PER_QUERY = {
    "kd_nodes": 85.531, "walk_tests": 5.919, "insphere": 44.259, "cavity": 20.311,
    "boundary": 27.089, "face_cc": 54.167, "sub3": 81.245, "sub4": 81.267, "nnz": 15.545,
    "tets_read": 44.261, "verts_read": 28.989,
}


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def plain_at_scale(n_keys=1_000_000, n_queries=10_000_000, seed=1):
    """Keys, queries and values of the plain-at-scale workload (seeded)."""
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


def bytes_per_query(w):
    """Algorithmic HBM bytes of the row kernel per query: the query (24 B),
    each k-d node entered (centroid + children, 32 B), each tet record read
    (4 vertex + 4 neighbour indices, 32 B), each distinct vertex read (24 B),
    the row written (index + weight, 12 B per entry) and its offset/length
    (12 B)."""
    return (24 + 32 * w["kd_nodes"] + 32 * w["tets_read"] + 24 * w["verts_read"]
            + 12 * w["nnz"] + 12)


def fp64_ops_per_query(w):
    """Non-fused FP64 operations per query: k-d distance and rectangle tests
    (~19 per node), walk orient3d (23), insphere (77), tet circumcentres for
    cavity tets and virtual tets (79 each), face circumcentres (56 each:
    distinct faces plus two per new subcell), subcell volumes (89 each),
    normalisation (2 per entry)."""
    return (19 * w["kd_nodes"] + 23 * w["walk_tests"] + 77 * w["insphere"]
            + 79 * (w["cavity"] + w["boundary"]) + 56 * (w["face_cc"] + 2 * w["sub4"])
            + 89 * (w["sub3"] + w["sub4"]) + 2 * w["nnz"] + 1)


def cpu_threads():
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


class RefBuild(threading.Thread):
    """Builds the reference mesh for the CPU baseline while the GPU runs."""

    def __init__(self, kc):
        super().__init__(daemon=True)
        self.kc, self.mesh, self.err, self.seconds = kc, None, None, 0.0

    def run(self):
        try:
            from oracle import ref
            t = time.time()
            self.mesh = ref.RefMesh(self.kc)
            self.seconds = time.time() - t
        except Exception as e:  # reported, the baseline is then omitted
            self.err = repr(e)


def cpu_baseline(refbuild, qc, target_s):
    refbuild.join()
    if refbuild.mesh is None:
        log(f"cpu baseline unavailable: {refbuild.err}")
        return None
    th = cpu_threads()
    m = refbuild.mesh
    m.get_weights(qc[:2000], num_threads=th)  # first call packs the mesh
    n = min(len(qc), 100_000)
    t = time.perf_counter()
    m.get_weights(qc[:n], num_threads=th)
    rate = n / (time.perf_counter() - t)
    n = int(min(len(qc), max(n, rate * target_s)))
    times = []
    for _ in range(3):
        t = time.perf_counter()
        m.get_weights(qc[:n], num_threads=th)
        times.append(time.perf_counter() - t)
    med = sorted(times)[1]
    import platform
    return {"value": n / med, "unit": "queries/s", "cores": th, "kind": "reference",
            "sample": f"first {n:,} queries of the workload, median of 3 runs after a "
                      f"warm-up call, mesh built by the reference in {refbuild.seconds:.0f} s; "
                      f"host {platform.processor() or platform.machine()}"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--keys", type=int, default=1_000_000)
    ap.add_argument("--queries", type=int, default=10_000_000)
    ap.add_argument("--cpu-baseline", choices=["reference", "none"], default="reference")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the query path runs on a single MI355X "
              "(multi-GPU sharding is a design only); run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import natinterp3d_amd
    from natinterp3d_amd.natinterp3d import spmm_device

    log(f"workload: {args.keys:,} keys, {args.queries:,} queries")
    keys, queries, values = plain_at_scale(args.keys, args.queries)
    keys = np.ascontiguousarray(keys, np.float64)
    centroid = keys.mean(axis=0)
    kc = np.ascontiguousarray(keys - centroid)
    qc = np.ascontiguousarray(queries - centroid)

    refbuild = None
    if args.cpu_baseline == "reference":
        from oracle import ref
        if ref.available():
            refbuild = RefBuild(kc)
            refbuild.start()
        else:
            log("oracle/_ref not built: no cpu baseline")

    t = time.time()
    mesh = natinterp3d_amd.MeshAndVerticesHIP(kc, device=0)
    build_s = time.time() - t
    info = mesh.info()
    log(f"mesh: {info['n_tets']:,} tets, k-d depth {info['kd_depth']}, built in {build_s:.1f} s")

    dev = torch.device("cuda", 0)
    qd = torch.from_numpy(qc).to(dev)
    xd = torch.from_numpy(values).to(dev)
    M = qc.shape[0]
    mesh.set_timing(True)

    def step():
        crow, col, val = mesh.weights_device(qd)
        return spmm_device(crow, col, val, xd, M), col.numel()

    for i in range(args.warmup):
        y, nnz = step()
        torch.cuda.synchronize()
        log(f"warmup {i + 1}/{args.warmup}: nnz {nnz:,}")

    rows_ms, rows_items, levels = [], [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        y, nnz = step()
        tm = mesh.timings()
        rows_ms.append(tm["rows_ms"][0])
        rows_items.append(tm["rows_items"][0])
        levels.append(len(tm["rows_ms"]))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    qps = M / (ms_per_step / 1e3)
    log(f"timed {args.steps} steps: {ms_per_step:.2f} ms/step, {qps / 1e6:.1f} M queries/s; "
        f"row kernel {np.mean(rows_ms):.2f} ms, levels {max(levels)}, last timings {tm}")

    bq = bytes_per_query(WORK_PER_QUERY)
    avg_rows_ms = float(np.mean(rows_ms))
    achieved = bq * float(np.mean(rows_items)) / (avg_rows_ms / 1e3) / 1e9
    ops = fp64_ops_per_query(WORK_PER_QUERY) * float(np.mean(rows_items)) / (avg_rows_ms / 1e3)

    base = None
    if refbuild is not None:
        log("cpu baseline: reference on the host cores")
        base = cpu_baseline(refbuild, qc, args.cpu_seconds)
        if base:
            log(f"cpu baseline {base['value'] / 1e6:.3f} M queries/s on {base['cores']} threads")

    out = {
        "metric": "Sibson queries/s fwd+bwd @1M keys vs plain GPU fwd; neighbor sets bit-exact",
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "plain at scale: 1,000,000 keys (90% uniform [0,1]^3, 10% in 20 "
                        "Gaussian blobs sigma 0.01), 10,000,000 queries uniform "
                        "[0.02,0.98]^3, values float32 [N,3]",
            "keys": args.keys, "queries": M, "nnz": int(nnz),
            "pass": "forward: all weight rows (FP64, bit-exact CSR in HBM) + W @ values",
            "mesh_build_s": round(build_s, 1),
        },
        "roofline": {
            "kernel": "sibson_rows_kernel",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "bytes_per_query": bq, "launch_ms": avg_rows_ms,
            "fp64_ops_per_s": ops, "fp64_frac": ops / FP64_OPS_PEAK,
        },
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
