"""The opt-in GPU Delaunay build (nat_mesh_build_ex with NAT_BUILD_GPU).

CPU: the build's per-thread steps (natinterp3d_amd/csrc/delaunay.h) run
sequentially by tests/emu/delaunay_emu.cpp, compiled here with the address
sanitizer: on the uniform, blob and jittered-lattice recipes the tets equal
the reference mesh's (the host builder, itself pinned to the compiled
reference), every face passes the exact empty-circumsphere test, links are
mutual and orientations positive; a small candidate budget per round gives
the same tets; repeated points are reported, not perturbed.

GPU: the same on the device, plus the rows.  The tets' order, their vertex
order and the k-d tree differ from the reference's, so the parity rule is
per row: sorted column sets equal to the oracle's on the reference mesh,
weights within 1e-12; over the GPU-built mesh itself the rows are
bit-identical to the oracle's on that mesh.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from natinterp3d_amd import workloads

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
W_TOL = 1e-12


def centred(keys, queries=None):
    keys = np.ascontiguousarray(keys, np.float64)
    c = keys.mean(axis=0)
    kc = np.ascontiguousarray(keys - c)
    return (kc, None) if queries is None else (kc, np.ascontiguousarray(queries - c))


def tet_set(tv):
    tv = np.asarray(tv)
    tv = tv[tv[:, 0] >= 0]
    return set(map(tuple, np.sort(tv, axis=1).tolist()))


# ------------------------------------------------------------ CPU emulation

@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("emu") / "delaunay_emu")
    clang = "/opt/rocm/llvm/bin/clang++"
    if not os.path.exists(clang):
        pytest.skip("clang++ of the ROCm toolchain not found")
    cmd = [clang, "-std=c++17", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I", os.path.join(REPO, "natinterp3d_amd", "csrc"),
           os.path.join(HERE, "emu", "delaunay_emu.cpp"),
           os.path.join(REPO, "natinterp3d_amd", "csrc", "mesh_builder.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def run_emu(exe, pts, tmp_path, slots=16384):
    pts = np.ascontiguousarray(pts, np.float64)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        f.write(struct.pack("i", len(pts)))
        f.write(pts.tobytes())
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0")
    r = subprocess.run([exe, fin, fout, str(slots)], capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    b = open(fout, "rb").read()
    rc, err, rounds, T = struct.unpack("4i", b[:16])
    tv = np.frombuffer(b[16:16 + 16 * T], np.int32).reshape(T, 4)
    tn = np.frombuffer(b[16 + 16 * T:16 + 32 * T], np.int32).reshape(T, 4)
    audit = np.frombuffer(b[16 + 32 * T:16 + 32 * T + 32], np.int64)
    return dict(rc=rc, err=err, rounds=rounds, tv=tv, tn=tn, audit=audit)


def recipe(name, n):
    if name == "uniform":
        return workloads.uniform_small(n_keys=n, n_queries=1)[0]
    if name == "blobs":
        return workloads.plain_at_scale(n_keys=n, n_queries=1)[0]
    side = int(round(n ** (1 / 3)))
    return workloads.lattice_jitter(n_side=side, n_queries=1)[0]


@pytest.mark.parametrize("name", ["uniform", "blobs", "lattice_jitter"])
def test_emulated_build_equals_reference_mesh(emu, tmp_path, name):
    """The reference moves a point whose FP64 test came out exactly 0 (on the
    jittered lattice, one of 2,197 by 3.4e-5); the GPU build keeps points in
    place, so it is given the reference mesh's final coordinates."""
    from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP
    kc, _ = centred(recipe(name, 2000))
    ref = MeshAndVerticesHIP(kc, device=-1).export()
    pts = np.ascontiguousarray(ref["verts"][:len(kc)])
    res = run_emu(emu, pts, tmp_path)
    assert res["rc"] == 0 and res["err"] == 0, res
    faces, nonpos, links, bad = res["audit"].tolist()
    assert faces > 0 and nonpos == 0 and links == 0 and bad == 0
    assert tet_set(res["tv"]) == tet_set(ref["tet_v"])


def test_emulated_build_with_few_candidates_per_round(emu, tmp_path):
    """At most 7 candidates per round: many more rounds, the same tets."""
    from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP
    kc, _ = centred(recipe("uniform", 600))
    res = run_emu(emu, kc, tmp_path, slots=7)
    assert res["rc"] == 0 and res["audit"][1:].sum() == 0
    assert tet_set(res["tv"]) == tet_set(MeshAndVerticesHIP(kc, device=-1).export()["tet_v"])


def test_emulated_build_reports_repeated_points(emu, tmp_path):
    kc, _ = centred(recipe("uniform", 500))
    kc = np.vstack([kc, kc[17]])
    res = run_emu(emu, kc, tmp_path)
    assert res["rc"] == -3 and res["err"] & 2


def test_emulated_build_exact_lattice(emu, tmp_path):
    """Co-spherical and coplanar points without jitter: either a valid
    Delaunay tetrahedralisation (exact tests, strict insphere) or a reported
    degeneracy -- never a broken mesh."""
    g = np.arange(5) / 4.0
    kc, _ = centred(np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3))
    res = run_emu(emu, kc, tmp_path)
    if res["rc"] == 0:
        assert res["audit"][1:].sum() == 0
    else:
        assert res["rc"] == -3 and res["err"] & (2 | 4)


# ---------------------------------------------------------------------- GPU

def sorted_rows(W):
    W = W.copy()
    W.sort_indices()
    return W


def rows_match(Wg, Wo, what):
    A, B = sorted_rows(Wg), sorted_rows(Wo)
    assert np.array_equal(A.indptr, B.indptr), f"{what}: row lengths differ"
    assert np.array_equal(A.indices, B.indices), f"{what}: column sets differ"
    err = float(np.abs(A.data - B.data).max()) if A.nnz else 0.0
    assert err <= W_TOL, f"{what}: max |dw| {err}"
    return err


def gpu_case(oracle_mod, keys, queries, label):
    from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP
    kc, qc = centred(keys, queries)
    print(f"[{label}] GPU build N={len(kc)}", flush=True)
    mg = MeshAndVerticesHIP(kc, device=0, builder="gpu")
    inf = mg.info()
    assert inf["builder"] == 1
    audit = mg.delaunay_audit()
    print(f"[{label}] {inf['n_tets']:,} tets in {inf['build_rounds']} rounds, "
          f"{inf['build_seconds']:.2f} s; audit {audit}", flush=True)
    assert audit["faces"] > 0 and audit["non_positive"] == 0
    assert audit["broken_links"] == 0 and audit["not_delaunay"] == 0
    mh = MeshAndVerticesHIP(kc, device=0)
    pg, ph = mg.export(), mh.export()
    assert np.array_equal(pg["verts"], ph["verts"])
    assert tet_set(pg["tet_v"]) == tet_set(ph["tet_v"])
    Wg = mg.get_weights(qc)
    Wo = oracle_mod.OracleMesh(ph).get_weights(qc)
    err = rows_match(Wg, Wo, label)
    Wself = oracle_mod.OracleMesh(pg).get_weights(qc)
    assert (np.array_equal(Wg.indptr, Wself.indptr) and np.array_equal(Wg.indices, Wself.indices)
            and np.array_equal(Wg.data, Wself.data)), f"{label}: rows over the GPU mesh"
    print(f"[{label}] rows: sorted columns equal, max |dw| {err:.3g}; bit-exact over the "
          f"GPU-built mesh", flush=True)


@pytest.mark.gpu
def test_gpu_build_uniform_10k(oracle_mod):
    """10,000 uniform keys, 100,000 queries in [0.05, 0.95]^3."""
    keys, queries, _ = workloads.uniform_small()
    gpu_case(oracle_mod, keys, queries, "gpu build uniform 10k/100k")


@pytest.mark.gpu
def test_gpu_build_plain_at_scale_recipe(oracle_mod):
    """100,000 keys of the plain-at-scale recipe (blobs), 1,000,000 queries."""
    keys, queries, _ = workloads.plain_at_scale(n_keys=100_000, n_queries=1_000_000)
    gpu_case(oracle_mod, keys, queries, "gpu build blobs 100k/1M")


@pytest.mark.gpu
def test_gpu_build_1m_keys():
    """1,000,000 keys of the plain-at-scale recipe: the build in seconds, an
    exactly Delaunay result."""
    from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP
    keys, queries, _ = workloads.plain_at_scale(n_keys=1_000_000, n_queries=10_000)
    kc, qc = centred(keys, queries)
    import time
    t0 = time.perf_counter()
    mg = MeshAndVerticesHIP(kc, device=0, builder="gpu")
    total = time.perf_counter() - t0
    inf = mg.info()
    audit = mg.delaunay_audit()
    print(f"[gpu build 1M] {inf['n_tets']:,} tets, {inf['build_rounds']} rounds, device build "
          f"{inf['build_seconds']:.2f} s, with packing and upload {total:.2f} s; audit {audit}",
          flush=True)
    assert audit["non_positive"] == 0 and audit["broken_links"] == 0
    assert audit["not_delaunay"] == 0
    W = mg.get_weights(qc)
    assert np.allclose(np.asarray(W.sum(axis=1)).ravel(), 1.0, atol=1e-12)


@pytest.mark.gpu
def test_gpu_build_rejects_repeated_points():
    from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP
    kc, _ = centred(recipe("uniform", 3000))
    kc = np.ascontiguousarray(np.vstack([kc, kc[5]]))
    with pytest.raises(RuntimeError, match="repeated"):
        MeshAndVerticesHIP(kc, device=0, builder="gpu")


@pytest.mark.gpu
def test_delaunay_audit_catches_a_planted_violation():
    """A host mesh whose vertex 0 is moved far off its place is no longer
    Delaunay (and some tets flip): the audit says so."""
    from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP
    kc, _ = centred(recipe("uniform", 3000))
    m = MeshAndVerticesHIP(kc, device=0)
    a = m.delaunay_audit()
    assert a["not_delaunay"] == 0 and a["non_positive"] == 0 and a["broken_links"] == 0
    pk = m.export()
    pk["verts"] = pk["verts"].copy()
    pk["verts"][0] += 0.05
    bad = MeshAndVerticesHIP.from_packed(pk, device=0).delaunay_audit()
    assert bad["not_delaunay"] > 0 or bad["non_positive"] > 0
