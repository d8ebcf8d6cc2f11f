"""ctypes access to oracle/_ref/libnatref.so -- the reference's own C sources
compiled where they lie (see oracle/Makefile and oracle/ref_glue.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke`` and by
``bench.py``'s CPU-baseline leg, never by the product package.
"""
import ctypes
import os

import numpy as np
import scipy.sparse

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(_HERE, "_ref", "libnatref.so")

_lib = None


def available():
    return os.path.exists(REF_SO)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError(f"{REF_SO} not built (needs the reference sources)")
        L = ctypes.CDLL(REF_SO)
        vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        L.ref_build.restype = vp
        L.ref_build.argtypes = [dp, ctypes.c_int, ctypes.c_int]
        L.ref_free.argtypes = [vp]
        L.ref_num_tets.restype = ctypes.c_int
        L.ref_num_tets.argtypes = [vp]
        L.ref_degeneracies.argtypes = [vp, ip, ip]
        L.ref_export.restype = ctypes.c_int
        L.ref_export.argtypes = [vp, dp, ip, ip, ip, ip, dp, ip, dp]
        L.ref_get_weights.restype = ctypes.c_int
        L.ref_get_weights.argtypes = [vp, dp, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(dp), ctypes.POINTER(ip), ip]
        L.ref_free_buf.argtypes = [vp]
        L.ref_srand.argtypes = [ctypes.c_uint]
        L.ref_rand.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class RefMesh:
    """The reference mesh (natural.c:114-118) built from already-centred keys."""

    def __init__(self, centered_keys, reset_rng=True):
        k = np.ascontiguousarray(centered_keys, np.float64)
        assert k.ndim == 2 and k.shape[1] == 3
        self.n = k.shape[0]
        self._h = lib().ref_build(_dp(k), self.n, 1 if reset_rng else 0)
        self._exported = None

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ref_free(self._h)
            self._h = None

    @property
    def num_tets(self):
        return lib().ref_num_tets(self._h)

    def degeneracies(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        lib().ref_degeneracies(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def export(self):
        """Packed arrays in list order; must run before the first get_weights."""
        if self._exported is not None:
            return self._exported
        T, n = self.num_tets, self.n
        out = dict(
            verts=np.zeros((n + 4, 3), np.float64),
            tet_v=np.zeros((T, 4), np.int32),
            tet_n=np.zeros((T, 4), np.int32),
            kd_left=np.zeros(T, np.int32),
            kd_right=np.zeros(T, np.int32),
            kd_pos=np.zeros((T, 3), np.float64),
            kd_data=np.zeros(T, np.int32),
            kd_rect=np.zeros(6, np.float64),
        )
        rc = lib().ref_export(self._h, _dp(out["verts"]), _ip(out["tet_v"]), _ip(out["tet_n"]),
                              _ip(out["kd_left"]), _ip(out["kd_right"]), _dp(out["kd_pos"]),
                              _ip(out["kd_data"]), _dp(out["kd_rect"]))
        if rc != 0:
            raise RuntimeError(f"ref_export failed: {rc}")
        self._exported = out
        return out

    def get_weights(self, centered_queries, num_threads=0):
        """Reference CSR (natural_insertionfree.c:761-940); num_threads<=0 = serial."""
        q = np.ascontiguousarray(centered_queries, np.float64)
        M = q.shape[0]
        rowptr = np.empty(M + 1, np.int32)
        vals = ctypes.POINTER(ctypes.c_double)()
        cols = ctypes.POINTER(ctypes.c_int)()
        rc = lib().ref_get_weights(self._h, _dp(q), M, int(num_threads),
                                   ctypes.byref(vals), ctypes.byref(cols), _ip(rowptr))
        if rc != 0:
            raise OverflowError("Total number of nonzero weights exceeds int32 limit")
        nnz = int(rowptr[M])
        v = np.ctypeslib.as_array(vals, shape=(nnz,)).copy() if nnz else np.zeros(0, np.float64)
        c = np.ctypeslib.as_array(cols, shape=(nnz,)).copy() if nnz else np.zeros(0, np.int32)
        lib().ref_free_buf(ctypes.cast(vals, ctypes.c_void_p))
        lib().ref_free_buf(ctypes.cast(cols, ctypes.c_void_p))
        return scipy.sparse.csr_matrix((v, c, rowptr), shape=(M, self.n))


def libc_rand_sequence(seed, count):
    """`count` outputs of the libc rand() the reference perturbs with."""
    L = lib()
    L.ref_srand(seed)
    return np.array([L.ref_rand() for _ in range(count)], np.int64)
