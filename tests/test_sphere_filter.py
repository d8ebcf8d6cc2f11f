"""The cavity search's certified in-sphere shortcut (csrc/sphere.h) against
the plain FP64 insphere determinant (predicates.c:3301-3347 as restated in
csrc/geom.h): wherever the shortcut certifies a side, it must be the plain
determinant's, on well-shaped tets, slivers, near-cospherical lattice tets,
the builder's own tets (super-simplex tets included) and queries placed on,
just inside and just outside the circumspheres.  Runs on the CPU
(tests/emu/sphere_emu.cpp, the header compiled for the host).

A planted mistake -- the shortcut without its orientation condition --
must be caught by the same kind of inputs.
"""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CLANG = "/opt/rocm/llvm/bin/clang++"


def _build(tmp, extra=()):
    exe = str(tmp / "sphere_emu")
    if not os.path.exists(CLANG):
        pytest.skip("clang++ of the ROCm toolchain not found")
    cmd = [CLANG, "-std=c++17", "-O2", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", "-I", os.path.join(REPO, "natinterp3d_amd", "csrc"),
           *extra, os.path.join(HERE, "emu", "sphere_emu.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    return _build(tmp_path_factory.mktemp("sphere"))


def run(exe, tets, queries, tmp):
    tets = np.ascontiguousarray(tets, np.float64).reshape(-1, 4, 3)
    K = len(tets)
    queries = np.ascontiguousarray(queries, np.float64).reshape(K, -1, 3)
    Q = queries.shape[1]
    fin, fout = str(tmp / "in.bin"), str(tmp / "out.bin")
    with open(fin, "wb") as fh:
        fh.write(np.array([K, Q], np.int32).tobytes())
        fh.write(tets.tobytes())
        fh.write(queries.tobytes())
    subprocess.run([exe, fin, fout], check=True)
    raw = np.fromfile(fout, np.int8)
    return raw[:K * Q].reshape(K, Q), raw[K * Q:].reshape(K, Q)


def circumspheres(tets):
    """numpy float64 circumcentres and radii (to place queries only)"""
    a = tets[:, 0]
    u = tets[:, 1:] - a[:, None]
    rhs = 0.5 * (u ** 2).sum(-1)
    c = np.linalg.solve(u, rhs[..., None])[..., 0]
    return a + c, np.linalg.norm(c, axis=1)


def sphere_queries(rng, tets, per=48):
    """points on, just inside and just outside each circumsphere, at random
    points near the tet, and the tet's own vertices (exactly cospherical)"""
    c, r = circumspheres(tets)
    K = len(tets)
    d = rng.standard_normal((K, per, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    rel = np.array([0.0, 1e-16, -1e-16, 1e-14, -1e-14, 1e-12, -1e-12, 1e-10, -1e-10, 1e-8,
                    -1e-8, 1e-6, -1e-6, 1e-3, -1e-3, 0.3, -0.3, 2.0])
    f = rel[rng.integers(0, len(rel), (K, per))]
    q = c[:, None] + (r[:, None, None] * (1.0 + f[..., None])) * d
    q[:, :4] = tets
    q[:, 4:8] = tets.mean(axis=1)[:, None] + 0.5 * (tets - tets.mean(axis=1)[:, None])
    return q


def check(fast, plain, what, min_rate=None):
    cert = fast >= 0
    bad = cert & (fast != plain)
    assert not bad.any(), f"{what}: {int(bad.sum())} certified decisions differ from FP64"
    if min_rate is not None:
        assert cert.mean() >= min_rate, f"{what}: only {cert.mean():.3f} certified"
    return cert.mean()


def test_random_tets(emu, tmp_path):
    rng = np.random.default_rng(0)
    tets = rng.random((20000, 4, 3))
    fast, plain = run(emu, tets, sphere_queries(rng, tets), tmp_path)
    rate = check(fast, plain, "random tets")
    # half the queries lie within 1e-10 of the sphere (relative), where the
    # shortcut gives way to the determinant; the rest are certified
    assert rate > 0.45


def test_slivers_and_caps(emu, tmp_path):
    rng = np.random.default_rng(1)
    K = 20000
    # four points on a sphere near one great circle (slivers), and four
    # points on a small cap (flat tets with huge circumspheres)
    th = rng.random((K, 4)) * 2 * np.pi
    z = rng.uniform(-1, 1, (K, 1)) * 10.0 ** rng.uniform(-12, -1, (K, 1))
    sl = np.stack([np.cos(th), np.sin(th), z * rng.standard_normal((K, 4))], axis=2)
    cap = rng.standard_normal((K, 4, 3)) * 10.0 ** rng.uniform(-9, -2, (K, 1, 1))
    cap[..., 2] = 0.5 * (cap[..., 0] ** 2 + cap[..., 1] ** 2)
    tets = np.concatenate([sl, cap])
    ok = np.abs(np.linalg.det(tets[:, 1:] - tets[:, :1])) > 0
    tets = tets[ok]
    fast, plain = run(emu, tets, sphere_queries(rng, tets), tmp_path)
    check(fast, plain, "slivers and caps")


def test_jittered_lattice_tets(emu, tmp_path):
    """the builder's tets of a jittered 12^3 lattice (near-cospherical: the 8
    corners of every cell lie within 1e-9 of one sphere), queried at the
    lattice points around them and near their spheres"""
    from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP
    rng = np.random.default_rng(2)
    g = np.stack(np.meshgrid(*[np.arange(12)] * 3, indexing="ij"), -1).reshape(-1, 3) / 11.0
    keys = g + rng.uniform(-1e-9, 1e-9, g.shape)
    keys -= keys.mean(axis=0)
    p = MeshAndVerticesHIP(np.ascontiguousarray(keys), device=-1).export()
    V = p["verts"].reshape(-1, 3)
    tv = p["tet_v"].reshape(-1, 4)
    tets = V[tv]
    data = (tv < len(keys)).all(axis=1)
    T = tets[data]
    # every lattice point within two spacings of the tet's centroid
    cen = T.mean(axis=1)
    near = []
    for t in range(len(T)):
        dd = np.abs(keys - cen[t]).max(axis=1)
        idx = np.nonzero(dd < 2.2 / 11.0)[0][:48]
        qq = np.empty((48, 3))
        qq[:] = keys[idx[0]]
        qq[:len(idx)] = keys[idx]
        near.append(qq)
    fast, plain = run(emu, T, np.stack(near), tmp_path)
    check(fast, plain, "lattice points")
    fast, plain = run(emu, T, sphere_queries(rng, T), tmp_path)
    check(fast, plain, "lattice spheres")
    # tets with a super vertex (huge, flat circumspheres near the hull)
    S = tets[~data]
    fast, plain = run(emu, S, sphere_queries(rng, S) * 0.0 + keys[rng.integers(0, len(keys),
                                                                               (len(S), 48))],
                      tmp_path)
    check(fast, plain, "super-simplex tets")


def test_mesh_tets_with_workload_queries(emu, tmp_path):
    """the builder's tets of blob-clustered keys, queried at uniform points
    and at the keys themselves; most decisions are certified"""
    from natinterp3d_amd import workloads
    from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP
    rng = np.random.default_rng(3)
    keys, queries, _ = workloads.plain_at_scale(n_keys=20000, n_queries=200000)
    keys = keys - keys.mean(axis=0)
    p = MeshAndVerticesHIP(np.ascontiguousarray(keys), device=-1).export()
    V = p["verts"].reshape(-1, 3)
    tv = p["tet_v"].reshape(-1, 4)
    T = V[tv]
    pick = rng.choice(len(T), 20000, replace=False)
    T = T[pick]
    c, r = circumspheres(T)
    q = c[:, None] + r[:, None, None] * rng.uniform(-1.5, 1.5, (len(T), 48, 3))
    q[:, :16] = keys[rng.integers(0, len(keys), (len(T), 16))]
    fast, plain = run(emu, T, q, tmp_path)
    check(fast, plain, "mesh tets", min_rate=0.95)


def test_scaled_and_offset_inputs(emu, tmp_path):
    rng = np.random.default_rng(4)
    base = rng.random((4000, 4, 3))
    for scale, off in ((1e-12, 0.0), (1e12, 0.0), (1.0, 1e6), (1e-6, 1e3), (1e20, 0.0),
                       (1e-30, 0.0)):
        tets = base * scale + off
        fast, plain = run(emu, tets, sphere_queries(rng, tets), tmp_path)
        check(fast, plain, f"scale {scale} offset {off}")


def test_planted_mistake_is_caught(tmp_path):
    """a shortcut without the orientation condition (iii) -- m's sign taken
    whenever it clears the rounding margins -- certifies wrong sides of
    slivers, and these inputs show it"""
    src = open(os.path.join(REPO, "natinterp3d_amd", "csrc", "sphere.h")).read()
    assert "m2 * s.w > s2 * s2 * sum" in src
    planted = tmp_path / "planted"
    planted.mkdir()
    (planted / "sphere.h").write_text(src.replace("m2 * s.w > s2 * s2 * sum", "true"))
    exe = str(tmp_path / "planted_emu")
    cmd = [CLANG, "-std=c++17", "-O2", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", "-I", str(planted), "-I",
           os.path.join(REPO, "natinterp3d_amd", "csrc"),
           os.path.join(HERE, "emu", "sphere_emu.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    rng = np.random.default_rng(5)
    th = rng.random((20000, 4)) * 2 * np.pi
    z = 1e-12 * rng.standard_normal((20000, 4))
    tets = np.stack([np.cos(th), np.sin(th), z], axis=2)
    fast, plain = run(exe, tets, sphere_queries(rng, tets), tmp_path)
    assert ((fast >= 0) & (fast != plain)).any(), "the planted mistake went unnoticed"
