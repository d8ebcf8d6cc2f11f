"""The C ABI library loads, exports every entry point include/*.h declares,
and its argument checks and host-only paths work without a GPU."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from natinterp3d_amd import _lib
from natinterp3d_amd.natinterp3d import MeshAndVerticesHIP

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(nat_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 9
    for s in syms:
        assert hasattr(L, s), s
    assert set(_lib.EXPORTS) <= set(syms)


def test_library_is_built_for_gfx950():
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_host_only_mesh_info_and_export():
    rng = np.random.default_rng(0)
    k = rng.random((500, 3))
    m = MeshAndVerticesHIP(np.ascontiguousarray(k - k.mean(0)), device=-1)
    inf = m.info()
    assert inf["n_data"] == 500 and inf["n_tets"] > 2000 and inf["device"] == -1
    e = m.export()
    assert e["tet_v"].shape == (inf["n_tets"], 4)
    # every interior face is shared both ways
    tn = e["tet_n"]
    for t in range(0, inf["n_tets"], 37):
        for nb in tn[t]:
            if nb >= 0:
                assert t in tn[nb]


def test_compute_on_host_only_mesh_is_an_error():
    rng = np.random.default_rng(1)
    k = rng.random((100, 3))
    m = MeshAndVerticesHIP(np.ascontiguousarray(k - k.mean(0)), device=-1)
    with pytest.raises(RuntimeError, match="host-only"):
        m.get_weights(np.zeros((3, 3)))


def test_argument_checks():
    L = _lib.lib()
    out = ctypes.c_void_p()
    pts = np.zeros((4, 3))
    dp = pts.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert L.nat_mesh_build(dp, 0, -1, ctypes.byref(out)) < 0
    assert "data point" in _lib.last_error()
    assert L.nat_mesh_build(dp, 4, -7, ctypes.byref(out)) < 0
    assert L.nat_spmm_csr(-1, 1, 1, None, None, None, 1, None, 1, None, None, 1, None) < 0
    assert L.nat_spmm_csr(3, 1, 1, None, None, None, 1, None, 0, None, None, 0, None) < 0
    order = ctypes.c_void_p()
    assert L.nat_weights_row_order(None, 3, ctypes.byref(order)) < 0
    assert L.nat_topology_target(None, None, 0, 0) < 0


def test_query_shape_checks():
    rng = np.random.default_rng(2)
    k = rng.random((50, 3))
    m = MeshAndVerticesHIP(np.ascontiguousarray(k - k.mean(0)), device=-1)
    with pytest.raises(ValueError, match="Query points must have shape"):
        m.get_weights(np.zeros((3, 2)))
    with pytest.raises(ValueError):
        m.get_weights(np.zeros((3, 3), np.float32))


def test_policy_argument_checks():
    from natinterp3d_amd.natinterp3d import get_row_policy, set_row_policy
    rng = np.random.default_rng(3)
    k = rng.random((60, 3))
    m = MeshAndVerticesHIP(np.ascontiguousarray(k - k.mean(0)), device=-1)
    assert get_row_policy(m.handle, 1) == {"budgets": [96], "first_level": 1,
                                           "work_sort": True, "topo_grid": 1, "geom_grid": 1,
                                           "wave_rows": True, "geom_lanes": 2}
    for kw, msg in (({"first_level": 3}, "first_level"), ({"topo_grid": 0}, "grid"),
                    ({"geom_lanes": 3}, "geom_lanes"), ({"geom_lanes": -1}, "geom_lanes"),
                    ({"geom_grid": 5}, "grid"), ({"budgets": [97]}, "budget"),
                    ({"first_level": 2}, "host-only")):
        with pytest.raises(RuntimeError, match=msg):
            set_row_policy(m.handle, 1, **kw)
    with pytest.raises(ValueError):
        set_row_policy(m.handle, 1, budgets=[1, 2])
    rep = _lib.TuneReport()
    assert _lib.lib().nat_mesh_autotune(m.handle, None, 0, None, 0, None,
                                        ctypes.byref(rep)) < 0
    assert "repeats" in _lib.last_error()


def test_mesh_round_trip_through_packed_arrays(tmp_path):
    """A host-only mesh saved to .npz and loaded back (no builder run) exports
    the same arrays; corrupted arrays are rejected."""
    rng = np.random.default_rng(4)
    k = rng.random((300, 3))
    m = MeshAndVerticesHIP(np.ascontiguousarray(k - k.mean(0)), device=-1)
    path = tmp_path / "mesh.npz"
    m.save(path)
    m2 = MeshAndVerticesHIP.load(path, device=-1)
    a, b = m.export(), m2.export()
    for key in a:
        assert np.array_equal(a[key], b[key]), key
    i1, i2 = m.info(), m2.info()
    for key in ("n_data", "n_tets", "kd_depth", "coplanar", "cospherical", "edge_anomalies"):
        assert i1[key] == i2[key], key
    bad = dict(a)
    bad["tet_v"] = a["tet_v"].copy()
    bad["tet_v"][5, 2] = a["verts"].shape[0]  # vertex slot out of range
    with pytest.raises(RuntimeError, match="vertex slot"):
        MeshAndVerticesHIP.from_packed(bad, device=-1)
    bad = dict(a)
    bad["kd_left"] = a["kd_left"].copy()
    node = int(np.flatnonzero(a["kd_left"] >= 0)[0])
    bad["kd_left"][node] = 0  # a cycle back to the root
    with pytest.raises(RuntimeError, match="k-d"):
        MeshAndVerticesHIP.from_packed(bad, device=-1)
    with pytest.raises(ValueError, match="not a saved"):
        np.savez(tmp_path / "other.npz", x=np.zeros(3))
        MeshAndVerticesHIP.load(tmp_path / "other.npz", device=-1)
