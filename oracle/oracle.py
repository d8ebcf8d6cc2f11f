"""ctypes access to oracle/liboracle.so, the C restatement of the Sibson query
path (oracle/sibson_oracle.c).

TEST INFRASTRUCTURE ONLY: this is the checker.  Imported by tests/, by
``__graft_entry__.smoke`` and by ``bench.py``'s CPU-baseline leg; the product
package never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np
import scipy.sparse

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle.so")

_lib = None


class OrcMesh(ctypes.Structure):
    _fields_ = [
        ("n_data", ctypes.c_int),
        ("n_tets", ctypes.c_int),
        ("verts", ctypes.POINTER(ctypes.c_double)),
        ("tet_v", ctypes.POINTER(ctypes.c_int)),
        ("tet_n", ctypes.POINTER(ctypes.c_int)),
        ("kd_left", ctypes.POINTER(ctypes.c_int)),
        ("kd_right", ctypes.POINTER(ctypes.c_int)),
        ("kd_pos", ctypes.POINTER(ctypes.c_double)),
        ("kd_rect", ctypes.POINTER(ctypes.c_double)),
    ]


COUNTER_FIELDS = ["queries", "kd_nodes", "walk_tests", "insphere", "cavity", "boundary",
                  "face_cc", "sub3", "sub4", "nnz", "dist_verts", "max_cavity", "max_visited",
                  "max_stack", "max_boundary", "max_nnz", "coincident", "empty_rows",
                  "tets_read", "verts_read"]


class OrcCounters(ctypes.Structure):
    _fields_ = [(f, ctypes.c_longlong) for f in COUNTER_FIELDS]


def build():
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
        dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        L.orc_get_weights.restype = ctypes.c_int
        L.orc_get_weights.argtypes = [ctypes.POINTER(OrcMesh), dp, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(dp), ctypes.POINTER(ip), ip,
                                      ctypes.POINTER(OrcCounters)]
        L.orc_locate.argtypes = [ctypes.POINTER(OrcMesh), dp, ctypes.c_int, ip]
        L.orc_cavity.restype = ctypes.c_int
        L.orc_cavity.argtypes = [ctypes.POINTER(OrcMesh), dp, ip, ctypes.c_int]
        L.orc_orient3d.restype = ctypes.c_double
        L.orc_orient3d.argtypes = [dp] * 4
        L.orc_insphere.restype = ctypes.c_double
        L.orc_insphere.argtypes = [dp] * 5
        L.orc_cc_tri.argtypes = [dp] * 4
        L.orc_cc_tet.argtypes = [dp] * 5
        L.orc_subcell.restype = ctypes.c_double
        L.orc_subcell.argtypes = [dp] * 8
        L.orc_free.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class OracleMesh:
    """The checker over a packed mesh (dict with verts, tet_v, tet_n, kd_left,
    kd_right, kd_pos, kd_rect -- the layout of RefMesh.export())."""

    def __init__(self, packed):
        self.a = {k: np.ascontiguousarray(packed[k]) for k in
                  ("verts", "tet_v", "tet_n", "kd_left", "kd_right", "kd_pos", "kd_rect")}
        for k in ("tet_v", "tet_n", "kd_left", "kd_right"):
            self.a[k] = np.ascontiguousarray(self.a[k], np.int32)
        for k in ("verts", "kd_pos", "kd_rect"):
            self.a[k] = np.ascontiguousarray(self.a[k], np.float64)
        self.n = self.a["verts"].shape[0] - 4
        self.s = OrcMesh(self.n, self.a["tet_v"].shape[0], _dp(self.a["verts"]),
                         _ip(self.a["tet_v"]), _ip(self.a["tet_n"]), _ip(self.a["kd_left"]),
                         _ip(self.a["kd_right"]), _dp(self.a["kd_pos"]), _dp(self.a["kd_rect"]))

    def get_weights(self, centered_queries, num_threads=0, counters=False):
        q = np.ascontiguousarray(centered_queries, np.float64).reshape(-1, 3)
        M = q.shape[0]
        rowptr = np.empty(M + 1, np.int32)
        vals = ctypes.POINTER(ctypes.c_double)()
        cols = ctypes.POINTER(ctypes.c_int)()
        ctr = OrcCounters()
        rc = lib().orc_get_weights(ctypes.byref(self.s), _dp(q), M, int(num_threads),
                                   ctypes.byref(vals), ctypes.byref(cols), _ip(rowptr),
                                   ctypes.byref(ctr))
        if rc != 0:
            raise OverflowError("Total number of nonzero weights exceeds int32 limit")
        nnz = int(rowptr[M])
        v = np.ctypeslib.as_array(vals, shape=(nnz,)).copy() if nnz else np.zeros(0, np.float64)
        c = np.ctypeslib.as_array(cols, shape=(nnz,)).copy() if nnz else np.zeros(0, np.int32)
        lib().orc_free(ctypes.cast(vals, ctypes.c_void_p))
        lib().orc_free(ctypes.cast(cols, ctypes.c_void_p))
        W = scipy.sparse.csr_matrix((v, c, rowptr), shape=(M, self.n))
        if counters:
            return W, {f: int(getattr(ctr, f)) for f in COUNTER_FIELDS}
        return W

    def locate(self, centered_queries):
        q = np.ascontiguousarray(centered_queries, np.float64).reshape(-1, 3)
        out = np.empty(q.shape[0], np.int32)
        lib().orc_locate(ctypes.byref(self.s), _dp(q), q.shape[0], _ip(out))
        return out

    def cavity(self, query):
        q = np.ascontiguousarray(query, np.float64).reshape(3)
        out = np.empty(4096, np.int32)
        n = lib().orc_cavity(ctypes.byref(self.s), _dp(q), _ip(out), 4096)
        return out[:min(n, 4096)].copy()


def orient3d(a, b, c, d):
    a, b, c, d = (np.ascontiguousarray(x, np.float64) for x in (a, b, c, d))
    return lib().orc_orient3d(_dp(a), _dp(b), _dp(c), _dp(d))


def insphere(a, b, c, d, e):
    a, b, c, d, e = (np.ascontiguousarray(x, np.float64) for x in (a, b, c, d, e))
    return lib().orc_insphere(_dp(a), _dp(b), _dp(c), _dp(d), _dp(e))


def cc_tri(a, b, c):
    a, b, c = (np.ascontiguousarray(x, np.float64) for x in (a, b, c))
    out = np.zeros(3)
    lib().orc_cc_tri(_dp(a), _dp(b), _dp(c), _dp(out))
    return out


def cc_tet(a, b, c, d):
    a, b, c, d = (np.ascontiguousarray(x, np.float64) for x in (a, b, c, d))
    out = np.zeros(3)
    lib().orc_cc_tet(_dp(a), _dp(b), _dp(c), _dp(d), _dp(out))
    return out


def subcell(pk, pa, pb, pc, ct, fab, fac, fbc):
    xs = [np.ascontiguousarray(x, np.float64) for x in (pk, pa, pb, pc, ct, fab, fac, fbc)]
    return lib().orc_subcell(*[_dp(x) for x in xs])
