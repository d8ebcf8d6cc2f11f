#!/usr/bin/env python3
"""Sibson query throughput, forward + backward, on one MI355X.

Workload (the plain-at-scale configuration): 1,000,000 keys (90% uniform in
[0,1]^3, 10% in 20 isotropic Gaussian blobs of sigma 0.01), 10,000,000
queries uniform in [0.02,0.98]^3, values float32 [N,3].  The mesh is built
once on the host (untimed, reported on stderr).  One step is one forward and
backward pass of the differentiable interpolation over all queries, inputs
resident in HBM: every weight row on the GPU (float64, bit-exact CSR in
device memory), y = W @ values, then dL/dvalues = W^T g and dL/dqueries
from the cached cavities.  The plain forward (rows + W @ values) is timed
beside it for the fwd+bwd / fwd ratio.

Prints one JSON line on stdout.  ``roofline`` is the dominant kernel's
achieved algorithmic bandwidth (bytes per query from the oracle's work
counters of this workload, times queries per launch, over the launch time
measured with HIP events on the launch stream) against MI355X's 8 TB/s.
``cpu_baseline`` is the compiled reference (oracle/_ref) on the host's cores
over a bounded sample of the same queries (forward only: the reference has no
backward).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import ctypes
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


def record_bytes(units, M):
    """Mean topology-record bytes per query (16-byte units) of the run."""
    return 16.0 * units / M


def kernel_bytes_per_query(w, D):
    """Algorithmic HBM bytes per query of each level-0 kernel, from the mean
    work counters w (cavity tets C, boundary faces B, insphere-tested tets T,
    distinct vertices V, row entries nnz, which is also the neighbour count):

    topo_search_kernel: order + start (8), query (24), vertex and neighbour
      ids of every tested tet (32 T), their vertices (24 V), the record it
      writes: header, C entries and the handoff (tet, adjacency) (16 + 24 C);
    topo_dfs_kernel + topo_order_kernel: the header and handoff read back
      (16 + 8 C), the cavity order (2 C) and boundary faces (B) written and
      read back, the cavity tets' vertex ids (16 C), the entries, neighbour
      list and header written (16 + 16 C + 4 nnz);
    geom_rows_kernel: order + record offset (12), query (24), the record
      sections it reads (header, entries, boundary faces, neighbour list:
      16 + 16 C + B + 4 nnz), the step-3 table's slice per cavity tet (32 C),
      tet vertex ids of cavity tets (16 C), vertices (24 V), row written
      (12 nnz) and its offset/length (12);
    grad_geom_kernel: as geom_rows_kernel without the row, plus dL/dy (8 D),
      the neighbours' values (8 D nnz) and dL/dq written (24)."""
    C, B, T, V, nnz = w["cavity"], w["boundary"], w["tets_read"], w["verts_read"], w["nnz"]
    rec_read = 16 + 16 * C + B + 4 * nnz
    geo = 12 + 24 + rec_read + 16 * C + 24 * V
    return {
        "topo_search_kernel": 8 + 24 + 32 * T + 24 * V + 16 + 24 * C,
        "topo_dfs_kernel+topo_order_kernel": (16 + 8 * C) + 2 * (2 * C + B) + 16 * C
                                             + (16 + 16 * C + 4 * nnz),
        "geom_rows_kernel": geo + 32 * C + 12 * nnz + 12,
        "grad_geom_kernel": geo + 8 * D + 8 * D * nnz + 24,
    }


def kernel_fp64_ops_per_query(w):
    """Non-fused FP64 operations per query of each level-0 kernel, counted
    as the reference's path does them: insphere 77; tet circumcentre 79
    (cavity and virtual tets); face circumcentre 56 (distinct faces plus two
    per new subcell); subcell volume 89; normalisation 2 per entry; facet
    moment ~90 per new subcell and the gradient sums ~15 per volume."""
    geo = (79 * (w["cavity"] + w["boundary"]) + 56 * (w["face_cc"] + 2 * w["sub4"])
           + 89 * (w["sub3"] + w["sub4"]))
    return {
        "topo_search_kernel": 77 * w["insphere"],
        "topo_dfs_kernel+topo_order_kernel": 0,
        "geom_rows_kernel": geo + 2 * w["nnz"] + 1,
        "grad_geom_kernel": geo + 90 * w["sub4"] + 15 * (w["sub3"] + w["sub4"]),
    }


def cpu_threads():
    """Host threads for the CPU baseline: the affinity count of this process
    (the cores it may run on), and OMP_NUM_THREADS when the environment sets
    one (reported beside it, not applied: the reference's OpenMP loop is
    given the thread count explicitly)."""
    n = max(1, len(os.sched_getaffinity(0)))
    env = os.environ.get("OMP_NUM_THREADS")
    return n, (int(env) if env and env.isdigit() else None)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def device_mesh(kc, path):
    """The device mesh of the centred keys kc: loaded from ``path`` when that
    file holds the mesh of the same keys (a digest of kc is stored with it),
    else built on the host (and saved to ``path`` when one is given).  Returns
    (mesh, seconds, "built" | "loaded")."""
    import hashlib

    import natinterp3d_amd
    digest = hashlib.blake2b(memoryview(np.ascontiguousarray(kc)).cast("B"),
                             digest_size=20).hexdigest()
    t = time.time()
    if path and os.path.exists(path):
        with np.load(path, allow_pickle=False) as z:
            same = "keys_digest" in z and str(z["keys_digest"]) == digest
        if same:
            mesh = natinterp3d_amd.MeshAndVerticesHIP.load(path, device=0)
            return mesh, time.time() - t, "loaded"
    mesh = natinterp3d_amd.MeshAndVerticesHIP(kc, device=0)
    build_s = time.time() - t
    if path:
        mesh.save(path, keys_digest=np.array(digest))
    return mesh, build_s, "built"


class RefBuild(threading.Thread):
    """Builds the reference mesh for the CPU baseline beside the product's
    own (host) mesh build; joined before any GPU warm-up, so that nothing
    runs on the host during the timed GPU steps."""

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


def join_refbuild(refbuild):
    if refbuild is not None:
        t = time.time()
        refbuild.join()
        log(f"reference mesh for the cpu baseline: built in {refbuild.seconds:.1f} s "
            f"(waited {time.time() - t:.1f} s before the GPU warm-up)")


def cpu_baseline(refbuild, qc, target_s, runs=5):
    """The compiled reference's OpenMP forward (getInsertionFreeWeightsParallel)
    on the host cores: first a warm-up call (it packs the mesh, as the
    reference's first call does), then the median of `runs` timed calls over
    the first n queries, n sized so that the runs take about target_s."""
    refbuild.join()
    if refbuild.mesh is None:
        log(f"cpu baseline unavailable: {refbuild.err}")
        return None
    th, omp = cpu_threads()
    m = refbuild.mesh
    m.get_weights(qc[:2000], num_threads=th)  # first call packs the mesh
    n = min(len(qc), 100_000)
    t = time.perf_counter()
    m.get_weights(qc[:n], num_threads=th)
    rate = n / (time.perf_counter() - t)
    n = int(min(len(qc), max(n, rate * target_s / runs)))
    times = []
    for _ in range(runs):
        t = time.perf_counter()
        m.get_weights(qc[:n], num_threads=th)
        times.append(time.perf_counter() - t)
    med = float(np.median(times))
    model = cpu_model()
    return {"value": n / med, "unit": "queries/s", "cores": th, "kind": "reference",
            "threads": th, "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads_env": omp, "cpu_model": model, "runs": runs,
            "run_seconds": [round(x, 3) for x in times],
            "sample": f"first {n:,} queries of the workload, median of {runs} runs after a "
                      f"warm-up call, {th} OpenMP threads (the process's affinity count), "
                      f"mesh built by the reference in {refbuild.seconds:.0f} s (before the "
                      f"GPU warm-up); host CPU: {model}"}


def run_precision(args):
    """Precision / degeneracy stress: keys on a 100^3 lattice scaled to
    [0,1]^3 with per-coordinate jitter U[-1e-9, 1e-9] (near-cospherical
    tets), queries uniform in [0.01,0.99]^3, scalar float64 values.  One step
    = the forward pass (all weight rows + W @ values), FP64 (bit-identical to
    the reference) and mixed precision; reports the mixed weights' error
    against the FP64 weights, the share recomputed in FP64 and the insphere
    decisions the FP64 static error bound does not certify."""
    import torch

    import natinterp3d_amd
    from natinterp3d_amd import workloads
    from natinterp3d_amd.natinterp3d import row_order, spmm_device

    keys, queries, values = workloads.lattice_jitter(n_queries=args.queries)
    keys = np.ascontiguousarray(keys, np.float64)
    c = keys.mean(axis=0)
    log(f"precision workload: {len(keys):,} lattice keys, {len(queries):,} queries")
    t = time.time()
    mesh = natinterp3d_amd.MeshAndVerticesHIP(np.ascontiguousarray(keys - c), device=0)
    build_s = time.time() - t
    log(f"mesh: {mesh.info()['n_tets']:,} tets, built in {build_s:.1f} s "
        f"({mesh.info()['coplanar']} coplanar, {mesh.info()['cospherical']} cospherical fixes)")
    dev = torch.device("cuda", 0)
    qd = torch.from_numpy(np.ascontiguousarray(queries - c)).to(dev)
    xd = torch.from_numpy(np.ascontiguousarray(values[:, None])).to(dev)
    M = qd.shape[0]
    mesh.set_timing(True)

    def fwd():
        crow, col, val = mesh.weights_device(qd)
        return spmm_device(crow, col, val, xd, M, order=row_order(mesh.handle, M)), col, val

    def timed():
        for _ in range(args.warmup):
            fwd()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = fwd()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.steps, res

    ms64, (y64, col64, w64) = timed()
    geom64 = mesh.timings()["geom_ms"]
    crow64 = mesh.weights_device(qd)[0].clone()
    mesh.set_precision("mixed")
    msmx, (ymx, colmx, wmx) = timed()
    tm = mesh.timings()
    mesh.set_precision("fp64")
    # robust predicates: exact signs for the tests the static bound leaves open
    mesh.set_predicates("robust")
    msrb, (yrb, colrb, wrb) = timed()
    tr = mesh.timings()
    crowrb = mesh.weights_device(qd)[0]
    mesh.set_predicates("fp64")
    if torch.equal(crowrb, crow64):
        diff = (colrb != col64) | (wrb != w64)
        rid = torch.searchsorted(crow64[1:].long(), torch.nonzero(diff).flatten(), right=True)
        rows_rb = int(torch.unique(rid).numel())
    else:
        rows_rb = int((torch.diff(crowrb) != torch.diff(crow64)).sum().item())
    log(f"robust predicates {msrb:.2f} ms/step: exact evaluations {tr['robust_exact']:,}, "
        f"changed decisions {tr['robust_changed']:,}, unresolved {tr['robust_unresolved']:,}, "
        f"rows differing from FP64 {rows_rb:,}; mixed error statement "
        f"{tm['mixed_err_max']:.3g} over {tm['mixed_err_rows']:,} sampled rows")
    same_cols = bool(torch.equal(col64, colmx))
    werr = float((wmx - w64).abs().max()) if w64.numel() else 0.0
    yerr = float((ymx - y64).abs().max()) if M else 0.0
    decisions, uncertain = mesh.predicate_audit(qd)
    share = tm["mixed_fp64"] / max(1, tm["mixed_fp32"] + tm["mixed_fp64"])
    log(f"fp64 {ms64:.2f} ms/step (geometry {geom64:.2f} ms), mixed {msmx:.2f} ms/step "
        f"(geometry {tm['geom_ms']:.2f} ms); neighbour sets equal: {same_cols}; max weight "
        f"error {werr:.3g}, max value error {yerr:.3g}; FP64 tets/faces {share:.4f}, rows "
        f"redone {tm['mixed_rows_fp64']:,}; insphere decisions {decisions:,}, uncertified "
        f"{uncertain:,}")
    out = {
        "metric": "Sibson queries/s, forward, jittered-lattice stress: FP64 path and mixed "
                  "path, mixed error vs FP64",
        "value": M / (ms64 / 1e3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "precision: 1,000,000 keys on a 100^3 lattice in [0,1]^3 with jitter "
                        "U[-1e-9,1e-9], 10,000,000 queries uniform [0.01,0.99]^3, float64 "
                        "scalar values",
            "queries": M, "nnz": int(col64.numel()),
            "mixed_ms_per_step": msmx,
            "mixed_queries_per_s": M / (msmx / 1e3),
            "geometry_ms_fp64": geom64, "geometry_ms_mixed": tm["geom_ms"],
            "neighbour_sets_equal": same_cols,
            "mixed_max_abs_weight_error": werr,
            "mixed_max_abs_value_error": yerr,
            "mixed_fp64_share_tets_faces": share,
            "mixed_rows_redone_fp64": tm["mixed_rows_fp64"],
            "insphere_decisions": decisions,
            "insphere_uncertified": uncertain,
            "mixed_error_statement": {"max_abs_weight_error": tm["mixed_err_max"],
                                      "rows_sampled": tm["mixed_err_rows"],
                                      "entries_sampled": tm["mixed_err_entries"]},
            "robust_ms_per_step": msrb,
            "robust_exact_evaluations": tr["robust_exact"],
            "robust_changed_decisions": tr["robust_changed"],
            "robust_unresolved": tr["robust_unresolved"],
            "robust_rows_differing_from_fp64": rows_rb,
            "mesh_build_s": round(build_s, 1),
        },
    }
    print(json.dumps(out), flush=True)


def run_shell(args):
    """Differentiation workload: 20,000 keys on a closed noisy shell (unit
    sphere radially perturbed by 0.1 x 8 random real spherical harmonics of
    degree <= 4, N(0, 0.01) thickness), 2,000,000 queries uniform inside 0.9x
    the shell radius, values float64 [N,3].  One step = forward (all weight
    rows + W @ values) + autograd backward to values and queries.  Hollow
    shells give every query a cavity of hundreds to thousands of tets, so
    these rows come from the thread-per-query levels.  After the timed steps:
    a central finite-difference check (h = 1e-6) of dL/dq on 1,000 random
    queries."""
    import torch

    import natinterp3d_amd
    from natinterp3d_amd import workloads
    from natinterp3d_amd.autograd import SibsonInterpolation

    keys, queries, values, _ = workloads.shell(n_keys=args.shell_keys,
                                               n_queries=args.shell_queries)
    log(f"shell workload: {len(keys):,} keys, {len(queries):,} queries")
    centroid = np.ascontiguousarray(keys, np.float64).mean(axis=0)
    kc = np.ascontiguousarray(keys - centroid)
    qc = np.ascontiguousarray(queries - centroid)
    refbuild = None
    if args.cpu_baseline == "reference":
        from oracle import ref
        if ref.available():
            refbuild = RefBuild(kc)
            refbuild.start()
    mesh, build_s, how = device_mesh(kc, args.mesh_file)
    log(f"mesh: {mesh.info()['n_tets']:,} tets, {how} in {build_s:.1f} s")
    join_refbuild(refbuild)
    dev = torch.device("cuda", 0)
    M = qc.shape[0]
    qd = torch.from_numpy(qc).to(dev).requires_grad_(True)
    xd = torch.from_numpy(np.ascontiguousarray(values)).to(dev).requires_grad_(True)
    gy = torch.randn((M, 3), dtype=torch.float64, device=dev,
                     generator=torch.Generator(device=dev).manual_seed(6))
    mesh.set_timing(True)

    def step():
        qd.grad = None
        xd.grad = None
        y = SibsonInterpolation.apply(mesh, qd, xd)
        y.backward(gy)

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        torch.cuda.synchronize()
        log(f"step {i + 1}/{args.steps} done")
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    tm = mesh.timings()
    fwd_ms = float(sum(tm["rows_ms"])) + tm["wave_ms"]
    bwd_ms = float(sum(tm["grad_ms"])) + tm["grad_wave_ms"]
    log(f"fwd+bwd {ms:.1f} ms/step ({M / ms * 1e3 / 1e6:.2f} M queries/s); "
        f"timings {tm}")

    # finite differences on 1,000 random queries
    rng = np.random.default_rng(7)
    sel = torch.as_tensor(rng.choice(M, 1000, replace=False), device=dev)
    g_sel = gy[sel]
    q0 = qd.detach()[sel].clone()
    vals = xd.detach()

    def loss(qq):
        with torch.no_grad():
            crow, col, val = mesh.weights_device(qq)
            from natinterp3d_amd.natinterp3d import spmm_device
            return (spmm_device(crow, col, val, vals, qq.shape[0]) * g_sel).sum(dim=1)

    qa = q0.clone().requires_grad_(True)
    (SibsonInterpolation.apply(mesh, qa, vals) * g_sel).sum().backward()
    ana = qa.grad.detach()
    fd = torch.empty_like(ana)
    h = 1e-6
    for k in range(3):
        e = torch.zeros_like(q0)
        e[:, k] = h
        fd[:, k] = (loss(q0 + e) - loss(q0 - e)) / (2 * h)
    scale = torch.maximum(ana.norm(dim=1), ana.norm(dim=1).median())
    rel = ((fd - ana).norm(dim=1) / scale).cpu().numpy()
    log(f"finite differences (h=1e-6, 1000 queries): relative error median "
        f"{np.median(rel):.2e}, 99th percentile {np.percentile(rel, 99):.2e}, max {rel.max():.2e}")

    base = None
    if refbuild is not None:
        log("cpu baseline: reference on the host cores (forward only)")
        base = cpu_baseline(refbuild, qc, args.cpu_seconds)
    out = {
        "metric": "Sibson queries/s fwd+bwd, 20k-key hollow shell, 2M queries (differentiation "
                  "workload)",
        "value": M / (ms / 1e3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"shell: {len(keys):,} keys on a noisy harmonic shell, {M:,} queries "
                        "inside 0.9x its radius, values float64 [N,3]; forward + backward",
            "fwd_kernels_ms": fwd_ms, "bwd_kernels_ms": bwd_ms,
            "rows_items": tm["rows_items"], "grad_items": tm["grad_items"],
            "fd_check": {"h": h, "queries": 1000, "rel_err_median": float(np.median(rel)),
                         "rel_err_p99": float(np.percentile(rel, 99)),
                         "rel_err_max": float(rel.max()),
                         "scale": "max(|grad|, median |grad|) per query"},
        },
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


def run_batch(args):
    """Batching workload: independent problems with log-uniform key counts in
    [500, 50,000] and query counts in [1,000, 200,000], points uniform in
    per-problem boxes of sides U[0.5, 2], values float32 [N, 8]; every
    --shell-every-th problem is a hollow shell instead (keys [300, 3,000],
    queries [1,000, 10,000]).  One step = the forward pass (all weight rows +
    W @ values) of every problem.  Timed: one BatchInterpolator call per step
    (one launch per kernel for all problems) under the default row policy,
    then under the autotuner's policy, then the tuned launch sequence replayed
    as a hipGraph; and one single-problem call per problem (sequential)."""
    from concurrent.futures import ThreadPoolExecutor

    import torch

    import natinterp3d_amd
    from natinterp3d_amd import workloads
    from natinterp3d_amd.natinterp3d import row_order, spmm_device

    probs = workloads.batch_problems(n_problems=args.problems, shell_every=args.shell_every)
    nk = sum(len(k) for k, _, _ in probs)
    nq = sum(len(q) for _, q, _ in probs)
    log(f"batch workload: {len(probs)} problems, {nk:,} keys, {nq:,} queries")
    aff, omp = cpu_threads()
    th = min(aff, omp or 16)  # host builder pool: the box's CPU share
    t = time.time()
    batch = natinterp3d_amd.BatchInterpolator([k for k, _, _ in probs], device=0, num_threads=th)
    build_s = time.time() - t
    log(f"batched mesh built in {build_s:.1f} s on {th} threads")
    cents = [np.ascontiguousarray(k).mean(axis=0) for k, _, _ in probs]

    def single(i):
        k = np.ascontiguousarray(probs[i][0])
        return natinterp3d_amd.MeshAndVerticesHIP(np.ascontiguousarray(k - cents[i]), device=0)

    t = time.time()
    with ThreadPoolExecutor(th) as ex:
        singles = list(ex.map(single, range(len(probs))))
    log(f"{len(singles)} single meshes built in {time.time() - t:.1f} s")

    dev = torch.device("cuda", 0)
    qs = [np.ascontiguousarray(q - c) for (_, q, _), c in zip(probs, cents)]
    off = np.zeros(len(probs) + 1, np.int64)
    off[1:] = np.cumsum([len(q) for q in qs])
    qcat = torch.from_numpy(np.concatenate(qs)).to(dev)
    vcat = torch.from_numpy(np.concatenate([v for _, _, v in probs])).to(dev)
    qd = [torch.from_numpy(q).to(dev) for q in qs]
    vd = [torch.from_numpy(v).to(dev) for _, _, v in probs]

    def batched():
        crow, col, val = batch.weights_device(qcat, off)
        return spmm_device(crow, col, val, vcat, int(off[-1]),
                           order=row_order(batch.handle, int(off[-1]))), col.numel()

    def sequential():
        n = 0
        outs = []
        for m, q, v in zip(singles, qd, vd):
            crow, col, val = m.weights_device(q)
            outs.append(spmm_device(crow, col, val, v, q.shape[0],
                                    order=row_order(m.handle, q.shape[0])))
            n += col.numel()
        return outs, n

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.steps, res

    d_ms, (yd, _) = timed(batched)
    s_ms, (ys, nnz_s) = timed(sequential)
    t = time.time()
    report = batch.autotune_device(qcat, off, repeats=2)
    tune_s = time.time() - t
    policy = batch.policy()
    routed = [p for p, b in enumerate(policy["budgets"]) if b == 0]
    log(f"autotuner: {tune_s:.1f} s, {report['candidates']} candidates, compute "
        f"{report['default_ms']:.2f} -> {report['tuned_ms']:.2f} ms; {len(routed)} problems "
        f"routed to thread-per-query rows, first level {policy['first_level']}, work sort "
        f"{policy['work_sort']}, grids x{policy['topo_grid']} / x{policy['geom_grid']}")
    b_ms, (yb, nnz_b) = timed(batched)
    graph = natinterp3d_amd.WeightsGraph(batch, qcat, off)

    def replayed():
        crow, col, val = graph.run()
        return spmm_device(crow, col, val, vcat, int(off[-1])), col.numel()

    g_ms, (yg, _) = timed(replayed)
    same = all(torch.equal(yb[off[i]:off[i + 1]], ys[i]) for i in range(len(probs)))
    same = same and torch.equal(yb, yd) and torch.equal(yb, yg)
    log(f"batched: default policy {d_ms:.2f} ms/step, tuned {b_ms:.2f}, tuned graph "
        f"{g_ms:.2f}; sequential {s_ms:.2f} ms/step, speed-up {s_ms / b_ms:.2f}x, "
        f"nnz {nnz_b:,} / {nnz_s:,}, products equal: {same}")
    out = {
        "metric": f"batched Sibson queries/s, forward, {len(probs)} independent problems in one "
                  f"set of launches (autotuned row policy) vs {len(probs)} sequential "
                  "single-problem calls",
        "value": nq / (b_ms / 1e3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"batch: {len(probs)} problems, keys log-uniform [500, 50000], "
                        "queries log-uniform [1000, 200000], uniform in boxes of sides "
                        "U[0.5, 2], values float32 [N, 8]"
                        + (f"; every {args.shell_every}th problem a hollow shell (keys "
                           "[300, 3000], queries [1000, 10000])" if args.shell_every else ""),
            "keys": nk, "queries": nq, "nnz": int(nnz_b),
            "default_policy_ms_per_step": d_ms,
            "graph_ms_per_step": g_ms,
            "graph_queries_per_s": nq / (g_ms / 1e3),
            "tuner": {"seconds": round(tune_s, 1), "candidates": report["candidates"],
                      "compute_default_ms": report["default_ms"],
                      "compute_tuned_ms": report["tuned_ms"],
                      "left_group_kernel": report["left_group"],
                      "routed_problems": len(routed),
                      "routed_queries": report["routed_queries"],
                      "first_level": policy["first_level"], "work_sort": policy["work_sort"],
                      "topo_grid": policy["topo_grid"], "geom_grid": policy["geom_grid"]},
            "sequential_ms_per_step": s_ms,
            "sequential_queries_per_s": nq / (s_ms / 1e3),
            "speedup_vs_sequential": s_ms / b_ms,
            "products_equal_to_sequential": bool(same),
            "batched_mesh_build_s": round(build_s, 1),
        },
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--keys", type=int, default=1_000_000)
    ap.add_argument("--queries", type=int, default=10_000_000)
    ap.add_argument("--cpu-baseline", choices=["reference", "none"], default="reference")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--workload", choices=["plain", "batch", "precision", "shell"],
                    default="plain",
                    help="plain: the fwd+bwd headline; batch: 256 independent problems in "
                         "one set of launches against 256 sequential single-problem calls; "
                         "precision: jittered-lattice stress, FP64 against mixed precision; "
                         "shell: fwd+bwd on a hollow shell (differentiation workload)")
    ap.add_argument("--shell-keys", type=int, default=20_000)
    ap.add_argument("--shell-queries", type=int, default=2_000_000)
    ap.add_argument("--mixed", choices=["on", "off"], default="on",
                    help="plain workload: also time the step with mixed-precision rows")
    ap.add_argument("--autotune", choices=["on", "off"], default="on",
                    help="plain workload: tune the row policy on the queries before timing")
    ap.add_argument("--end-to-end", choices=["on", "off"], default="on",
                    help="plain workload: also time the drop-in host call (host queries in, "
                         "scipy CSR out), reported beside the device-resident value")
    ap.add_argument("--problems", type=int, default=256)
    ap.add_argument("--shell-every", type=int, default=0,
                    help="batch workload: every k-th problem a hollow shell (0: none)")
    ap.add_argument("--mesh-file", default=None,
                    help="plain and shell workloads: load the device mesh from this .npz "
                         "when it holds the mesh of the same keys, else build it on the "
                         "host and save it there (the build is outside the timed region "
                         "either way)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the steps, copy this many MiB device to device once: a "
                         "kernel of known traffic that scales hardware byte counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the query path runs on a single MI355X "
              "(multi-GPU sharding is a design only); run it with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "precision":
        return run_precision(args)
    if args.workload == "shell":
        return run_shell(args)

    import torch

    import natinterp3d_amd
    from natinterp3d_amd import workloads
    from natinterp3d_amd.autograd import SibsonInterpolation
    from natinterp3d_amd.natinterp3d import weights_spmm

    log(f"workload: {args.keys:,} keys, {args.queries:,} queries")
    keys, queries, values = workloads.plain_at_scale(n_keys=args.keys, n_queries=args.queries)
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

    mesh, build_s, how = device_mesh(kc, args.mesh_file)
    info = mesh.info()
    log(f"mesh: {info['n_tets']:,} tets, k-d depth {info['kd_depth']}, {how} in {build_s:.1f} s")
    join_refbuild(refbuild)

    dev = torch.device("cuda", 0)
    M, D = qc.shape[0], values.shape[1]
    qd = torch.from_numpy(qc).to(dev).requires_grad_(True)
    xd = torch.from_numpy(values).to(dev).requires_grad_(True)
    gy = torch.randn((M, D), dtype=torch.float64, device=dev,
                     generator=torch.Generator(device=dev).manual_seed(5))
    tuned = None
    if args.autotune == "on":
        t = time.time()
        rep_t = mesh.autotune(qd.detach(), repeats=2)
        tuned = {"seconds": round(time.time() - t, 1), **rep_t, **mesh.policy()}
        log(f"autotuner: {tuned}")
    mesh.set_timing(True)

    def step():
        """one forward + backward pass over all queries"""
        qd.grad = None
        xd.grad = None
        y = SibsonInterpolation.apply(mesh, qd, xd)
        y.backward(gy)
        return y

    def plain_forward():
        with torch.no_grad():
            crow, col, val = mesh.weights_device(qd.detach())
            return weights_spmm(mesh.handle, M, xd.detach()), col.numel()

    for i in range(args.warmup):
        y = step()
        torch.cuda.synchronize()
        log(f"warmup {i + 1}/{args.warmup}: timings {mesh.timings()}")

    per_step = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        y = step()
        per_step.append(mesh.timings())  # the backward's calls synchronise the stream
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    qps = M / (ms_per_step / 1e3)
    assert torch.isfinite(qd.grad).all() and torch.isfinite(xd.grad).all()

    # the plain forward, for the fwd+bwd / fwd ratio
    for _ in range(max(1, args.warmup // 2)):
        plain_forward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        _, nnz = plain_forward()
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - t2) * 1e3 / args.steps
    log(f"timed {args.steps} steps: fwd+bwd {ms_per_step:.2f} ms/step "
        f"({qps / 1e6:.1f} M queries/s), plain fwd {fwd_ms:.2f} ms/step, "
        f"ratio {ms_per_step / fwd_ms:.3f}; last timings {per_step[-1]}")

    # the same step with the mixed-precision rows (neighbour sets unchanged),
    # against the plain FP64 forward
    # the drop-in host call of the reference's plugin (get_weights on host
    # queries -> scipy CSR): H2D of the queries, the kernels, D2H of the CSR
    # over PCIe and the CSR wrap.  Reported only; `value` is device-resident.
    end_to_end = None
    if args.end_to_end == "on":
        runs = []
        for _ in range(2):
            t4 = time.perf_counter()
            W = mesh.get_weights(qc)
            runs.append(time.perf_counter() - t4)
            nnz_host = W.nnz
            del W
        best = min(runs)
        end_to_end = {"call": "MeshAndVerticesHIP.get_weights(host queries) -> scipy CSR",
                      "seconds": best, "queries_per_s": M / best, "nnz": nnz_host,
                      "csr_bytes": nnz_host * 12 + (M + 1) * 4, "runs_s": runs}
        log(f"drop-in get_weights (host in, scipy CSR out): {best * 1e3:.1f} ms "
            f"({M / best / 1e6:.1f} M queries/s, {nnz_host:,} entries)")

    mixed = None
    if args.mixed == "on":
        mesh.set_precision("mixed")
        step()
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        mixed_ms = (time.perf_counter() - t3) * 1e3 / args.steps
        tm = mesh.timings()
        mesh.set_precision("fp64")
        mixed = {"fwd_bwd_ms": mixed_ms, "fwd_bwd_over_plain_fp64_fwd": mixed_ms / fwd_ms,
                 "fp32_tets_faces": tm["mixed_fp32"], "fp64_tets_faces": tm["mixed_fp64"],
                 "rows_redone_fp64": tm["mixed_rows_fp64"]}
        log(f"mixed precision: fwd+bwd {mixed_ms:.2f} ms/step, "
            f"{mixed_ms / fwd_ms:.3f} x the plain FP64 forward")

    if args.calib_copy_mib > 0:
        n = args.calib_copy_mib << 20
        src = torch.ones(n, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {n:,} bytes read and written")
        del src, dst

    nbytes = ctypes.c_int64()
    natinterp3d_amd._lib.lib().nat_topology_save(mesh.handle, None, 0, ctypes.byref(nbytes),
                                                 None)
    # the records follow a 64-byte header and the per-query arrays (offsets,
    # order, problem, start tet, row total: 28 bytes per query, aligned)
    rec_start = ((((64 + 8 * M + 12 * M) + 7) & ~7) + 8 * M + 15) & ~15
    units = (int(nbytes.value) - rec_start) / 16.0
    rec_q = record_bytes(units, M)
    kb = kernel_bytes_per_query(WORK_PER_QUERY, D)
    kops = kernel_fp64_ops_per_query(WORK_PER_QUERY)
    kms = {"topo_search_kernel": float(np.mean([t["topo_search_ms"] for t in per_step])),
           "topo_dfs_kernel+topo_order_kernel":
               float(np.mean([t["topo_order_ms"] for t in per_step])),
           "geom_rows_kernel": float(np.mean([t["geom_ms"] for t in per_step])),
           "grad_geom_kernel": float(np.mean([t["grad_ms"][0] for t in per_step]))}
    dom = max(kms, key=kms.get)
    launch_ms = kms[dom]
    achieved = kb[dom] * M / (launch_ms / 1e3) / 1e9
    ops = kops[dom] * M / (launch_ms / 1e3)
    log(f"kernels (ms): {kms}; record {rec_q:.0f} B/query; dominant {dom}: "
        f"{kb[dom]:.0f} B/query -> {achieved:.1f} GB/s, {ops / 1e12:.2f} T FP64 op/s")

    # measured HBM bytes per launch of the dominant kernel (rocprofv3 --pmc on
    # this workload, committed with the commands that made them)
    traffic, traffic_src = None, None
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                         "hbm_traffic_plain_at_scale_1M_10M.txt")
    if args.keys == 1_000_000 and args.queries == 10_000_000 and os.path.exists(tpath):
        with open(tpath) as fh:
            for line in fh:
                name, _, val = line.partition("=")
                if not line.startswith("#") and name.strip() == dom and val.strip():
                    traffic = float(val)
                    traffic_src = "profiles/hbm_traffic_plain_at_scale_1M_10M.txt"

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
            "pass": "forward (all weight rows, FP64 bit-exact CSR in HBM, W @ values) + "
                    "backward (dL/dvalues = W^T g, dL/dqueries from cached cavities)",
            "plain_fwd_ms": fwd_ms,
            "fwd_bwd_over_fwd": ms_per_step / fwd_ms,
            "kernel_ms": {**kms, "locate_kernel": per_step[-1]["locate_ms"],
                          "wave_rows_kernel_fwd": per_step[-1]["wave_ms"],
                          "wave_rows_kernel_bwd": per_step[-1]["grad_wave_ms"],
                          "thread_levels_fwd": float(sum(per_step[-1]["rows_ms"][1:])),
                          "thread_levels_bwd": float(sum(per_step[-1]["grad_ms"][1:])),
                          "emit": per_step[-1]["emit_ms"]},
            "rows_beyond_group_kernel": per_step[-1]["wave_items"],
            "mixed_precision": mixed,
            "drop_in_end_to_end": end_to_end,
            "mesh_build_s": round(build_s, 1) if how == "built" else None,
            "row_policy": tuned if tuned is not None else "default (autotuner off)",
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_source": traffic_src,
            "bytes_per_query": kb[dom], "queries_per_launch": M, "launch_ms": launch_ms,
            "fp64_ops_per_s": ops, "fp64_frac": ops / FP64_OPS_PEAK,
        },
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
