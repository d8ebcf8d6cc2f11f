"""Public API, mirroring natinterp3d's (natinterp3d.py:9-137).

``Interpolator``, ``interpolate`` and ``get_weights`` keep the reference's
signatures, centring, dtype rules and return types.  Underneath, the mesh is
built once on the host (exactly the reference's mesh) and uploaded to the GPU;
every weight row and every ``W @ values`` product is computed by the gfx950
kernels of libnatinterp3d_hip.so.  ``parallel`` and ``num_threads`` are
accepted for compatibility and do not change any result (the reference's
results do not depend on them either, tests/test_natinterp3d.py:347-373).
"""
import collections
import ctypes
import hashlib
import os
import weakref
from typing import Optional

import numpy as np
import scipy.sparse

from . import _lib

_DT_CODE = {np.dtype(np.float32): 0, np.dtype(np.float64): 1}


def _free_host(addr):
    _lib.lib().nat_free_host(ctypes.c_void_p(addr))


def _adopt(ptr, n, ctype, dtype):
    """Wrap n elements of library-malloc'd memory as a NumPy array that frees
    it when collected (the reference hands its buffers to NumPy the same way,
    natinterp3d_cython.pyx:37-52)."""
    addr = ctypes.cast(ptr, ctypes.c_void_p).value
    if n == 0:
        if addr:
            _free_host(addr)
        return np.zeros(0, dtype)
    buf = (ctype * n).from_address(addr)
    fin = weakref.finalize(buf, _free_host, addr)
    fin.atexit = False
    return np.frombuffer(buf, dtype=dtype)


def _default_device():
    env = os.environ.get("NATINTERP3D_DEVICE")
    if env is not None:
        return int(env)
    try:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.current_device()
    except ImportError:
        pass
    return 0


class MeshAndVerticesHIP:
    """GPU-backed plugin object with the contract of the reference's
    ``MeshAndVertices`` / ``MeshAndVerticesParallel``
    (natinterp3d_cython.pyx:54-134): built from centred float64 data points,
    ``get_weights(centred_queries)`` returns a scipy CSR matrix
    [num_queries, num_data_points] with float64 data and int32 indices whose
    columns are in natural-neighbour discovery order."""

    def __init__(self, dataPoints, device: Optional[int] = None, builder: str = "host"):
        """builder "host" (default): the reference's mesh, rebuilt exactly on
        the host.  builder "gpu": the Delaunay tetrahedralisation made on the
        device (nat_mesh_build_ex with NAT_BUILD_GPU) -- the same tets for
        points in general position, in another order, so rows agree with the
        reference's as per-row column sets; degenerate input raises instead
        of being perturbed."""
        dataPoints = _as_c_double_2d(dataPoints)
        self.numDataPoints = dataPoints.shape[0]
        if dataPoints.shape[1] != 3:
            raise ValueError("Data points must have shape (num_data_points, 3)")
        builders = {"host": 0, "gpu": 1}
        if builder not in builders:
            raise ValueError(f"builder must be one of {sorted(builders)}")
        L = _lib.lib()
        self._m = ctypes.c_void_p()
        self.device = _default_device() if device is None else int(device)
        rc = L.nat_mesh_build_ex(dataPoints.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 self.numDataPoints, self.device, builders[builder],
                                 ctypes.byref(self._m))
        _lib.check(rc, "nat_mesh_build_ex")
        self._finalizer = weakref.finalize(self, L.nat_mesh_free, self._m)

    @property
    def handle(self):
        return self._m

    def info(self):
        inf = _lib.MeshInfo()
        _lib.check(_lib.lib().nat_mesh_get_info(self._m, ctypes.byref(inf)), "nat_mesh_get_info")
        return {f: getattr(inf, f) for f, _ in inf._fields_}

    def set_timing(self, enable=True):
        """Record HIP events around this mesh's kernels (see timings())."""
        _lib.check(_lib.lib().nat_mesh_set_timing(self._m, 1 if enable else 0),
                   "nat_mesh_set_timing")

    def timings(self):
        """Kernel times (ms) of the last weights computation on this mesh."""
        t = _lib.Timings()
        _lib.check(_lib.lib().nat_mesh_get_timings(self._m, ctypes.byref(t)),
                   "nat_mesh_get_timings")
        n = t.levels
        return {"order_ms": t.order_ms, "locate_ms": t.locate_ms, "topo_ms": t.topo_ms,
                "geom_ms": t.geom_ms,
                "rows_ms": list(t.rows_ms)[:n],
                "rows_items": list(t.rows_items)[:n], "emit_ms": t.emit_ms,
                "grad_ms": list(t.grad_ms), "grad_items": list(t.grad_items),
                "mixed_fp32": t.mixed_fp32, "mixed_fp64": t.mixed_fp64,
                "mixed_rows_fp64": t.mixed_rows_fp64, "wave_ms": t.wave_ms,
                "wave_items": t.wave_items, "grad_wave_ms": t.grad_wave_ms,
                "grad_wave_items": t.grad_wave_items, "topo_search_ms": t.topo_search_ms,
                "topo_order_ms": t.topo_order_ms, "robust_exact": t.robust_exact,
                "robust_changed": t.robust_changed, "robust_unresolved": t.robust_unresolved,
                "mixed_err_max": t.mixed_err_max, "mixed_err_rows": t.mixed_err_rows,
                "mixed_err_entries": t.mixed_err_entries}

    def set_precision(self, mode="fp64", cond_threshold=0.02, cancel_threshold=20.0):
        """Row arithmetic: "fp64" (default; weights bit-identical to the
        reference) or "mixed" (FP32 circumcentres and volumes relative to the
        query; FP64 for tets whose condition estimate is below cond_threshold
        and for rows whose subcell volumes cancel by more than
        cancel_threshold; neighbour sets unchanged)."""
        modes = {"fp64": 0, "mixed": 1}
        if mode not in modes:
            raise ValueError(f"precision must be one of {sorted(modes)}")
        _lib.check(_lib.lib().nat_mesh_set_precision(self._m, modes[mode], cond_threshold,
                                                     cancel_threshold),
                   "nat_mesh_set_precision")

    def set_predicates(self, mode="fp64"):
        """Predicates of the query path: "fp64" (default; the reference's
        plain FP64 signs, rows bit-identical to it) or "robust" (every
        insphere / orient3d sign the static error bound leaves uncertain is
        decided by the exact determinant).  timings() counts the exact
        evaluations (robust_exact), the decisions that changed against FP64
        (robust_changed) and the tests too wide to evaluate exactly
        (robust_unresolved, FP64 sign kept)."""
        modes = {"fp64": 0, "robust": 1}
        if mode not in modes:
            raise ValueError(f"predicates must be one of {sorted(modes)}")
        _lib.check(_lib.lib().nat_mesh_set_predicates(self._m, modes[mode]),
                   "nat_mesh_set_predicates")

    def set_mixed_check(self, stride=1024):
        """Error statement of mixed rows: every stride-th query is made
        again in FP64 after each mixed compute and timings() reports the
        largest |w_mixed - w_fp64| over them (mixed_err_max, with
        mixed_err_rows / mixed_err_entries covered).  0 turns it off."""
        _lib.check(_lib.lib().nat_mesh_set_mixed_check(self._m, int(stride)),
                   "nat_mesh_set_mixed_check")

    def policy(self):
        """The row-path policy (see set_policy)."""
        return get_row_policy(self._m, 1)

    def set_policy(self, **kw):
        """Change the row-path policy: budgets (list of one visited-tet budget,
        0..96; 0 = skip the group kernel), first_level (1 or 2), work_sort
        (bool), topo_grid, geom_grid (1..4), wave_rows (bool), geom_lanes
        (2 or 1: FP64 geometry kernel lanes per row).  Results do not depend
        on it."""
        set_row_policy(self._m, 1, **kw)

    def autotune(self, queries, repeats=3):
        """Time row-path policies on centred float64 CUDA queries and keep the
        fastest on this mesh; returns the tuner's report (dict)."""
        return autotune_rows(self._m, queries, None, repeats)

    def delaunay_audit(self):
        """Exact audit of the tetrahedralisation on the device: dict of faces
        tested, tets of non-positive orientation, non-mutual neighbour links
        and faces violating the empty-circumsphere property."""
        c = (ctypes.c_int64 * 4)()
        _lib.check(_lib.lib().nat_mesh_delaunay_audit(self._m, c), "nat_mesh_delaunay_audit")
        return {"faces": c[0], "non_positive": c[1], "broken_links": c[2],
                "not_delaunay": c[3]}

    def predicate_audit(self, queries):
        """Re-test the insphere decisions behind the rows of centred float64
        CUDA queries under the FP64 static error bound: returns (decisions,
        uncertain) -- how many cavity / boundary-neighbour tests the rows
        rest on and how many of their signs FP64 does not certify."""
        import torch
        q = queries.detach().to(dtype=torch.float64).contiguous()
        self.weights_device(q)
        L = _lib.lib()
        nbytes = ctypes.c_int64()
        _lib.check(L.nat_topology_save(self._m, None, 0, ctypes.byref(nbytes), None),
                   "nat_topology_save")
        topo = torch.empty(int(nbytes.value), dtype=torch.uint8, device=q.device)
        stream = ctypes.c_void_p(torch.cuda.current_stream(q.device).cuda_stream)
        _lib.check(L.nat_topology_save(self._m, ctypes.c_void_p(topo.data_ptr()), topo.numel(),
                                       ctypes.byref(nbytes), stream), "nat_topology_save")
        dec, unc = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(L.nat_predicate_audit(self._m, ctypes.c_void_p(topo.data_ptr()),
                                         ctypes.c_void_p(q.data_ptr() if q.numel() else 0),
                                         q.shape[0], ctypes.byref(dec), ctypes.byref(unc),
                                         stream), "nat_predicate_audit")
        return int(dec.value), int(unc.value)

    def export(self):
        """The packed host mesh: verts [(N+4),3], tet_v/tet_n [T,4], k-d arrays."""
        inf = self.info()
        T, n = inf["n_tets"], inf["n_data"]
        out = dict(verts=np.zeros((n + 4, 3)), tet_v=np.zeros((T, 4), np.int32),
                   tet_n=np.zeros((T, 4), np.int32), kd_left=np.zeros(T, np.int32),
                   kd_right=np.zeros(T, np.int32), kd_pos=np.zeros((T, 3)),
                   kd_rect=np.zeros(6))
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))  # noqa: E731
        rc = _lib.lib().nat_mesh_export(self._m, dp(out["verts"]), ip(out["tet_v"]),
                                        ip(out["tet_n"]), ip(out["kd_left"]),
                                        ip(out["kd_right"]), dp(out["kd_pos"]),
                                        dp(out["kd_rect"]))
        _lib.check(rc, "nat_mesh_export")
        return out

    _FORMAT = "natinterp3d_amd.mesh"

    def save(self, path, **extra):
        """Write the packed mesh (the builder's output: final coordinates,
        tets, neighbours, k-d tree) to ``path`` as .npz (plain arrays, no
        pickles); ``load`` makes the device mesh from it without running the
        builder, and its rows equal this mesh's."""
        inf = self.info()
        np.savez(path, format=np.array(self._FORMAT), version=np.array(1),
                 degeneracies=np.array([inf["coplanar"], inf["cospherical"],
                                        inf["edge_anomalies"]], np.int32),
                 **self.export(), **extra)

    @classmethod
    def load(cls, path, device: Optional[int] = None):
        """The mesh saved by ``save`` at ``path``, uploaded to ``device``."""
        with np.load(path, allow_pickle=False) as z:
            if "format" not in z or str(z["format"]) != cls._FORMAT or int(z["version"]) != 1:
                raise ValueError(f"{path}: not a saved natinterp3d_amd mesh")
            arrays = {k: z[k] for k in ("verts", "tet_v", "tet_n", "kd_left", "kd_right",
                                        "kd_pos", "kd_rect", "degeneracies")}
        return cls.from_packed(arrays, device=device)

    @classmethod
    def from_packed(cls, arrays, device: Optional[int] = None):
        """A mesh from packed arrays of the layout ``export`` returns (plus an
        optional ``degeneracies`` [3]); the arrays are checked by the
        library."""
        verts = np.ascontiguousarray(arrays["verts"], np.float64).reshape(-1, 3)
        tet_v = np.ascontiguousarray(arrays["tet_v"], np.int32).reshape(-1, 4)
        tet_n = np.ascontiguousarray(arrays["tet_n"], np.int32).reshape(-1, 4)
        kd_left = np.ascontiguousarray(arrays["kd_left"], np.int32).reshape(-1)
        kd_right = np.ascontiguousarray(arrays["kd_right"], np.int32).reshape(-1)
        kd_pos = np.ascontiguousarray(arrays["kd_pos"], np.float64).reshape(-1, 3)
        kd_rect = np.ascontiguousarray(arrays["kd_rect"], np.float64).reshape(-1)
        T = tet_v.shape[0]
        if (tet_n.shape[0] != T or kd_left.shape[0] != T or kd_right.shape[0] != T
                or kd_pos.shape[0] != T or kd_rect.shape[0] != 6 or verts.shape[0] < 5):
            raise ValueError("packed mesh arrays have inconsistent shapes")
        deg = arrays.get("degeneracies")
        deg = None if deg is None else np.ascontiguousarray(deg, np.int32).reshape(3)
        self = cls.__new__(cls)
        self.numDataPoints = verts.shape[0] - 4
        self.device = _default_device() if device is None else int(device)
        self._m = ctypes.c_void_p()
        L = _lib.lib()
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        rc = L.nat_mesh_from_packed(vp(verts), self.numDataPoints, vp(tet_v), vp(tet_n), T,
                                    vp(kd_left), vp(kd_right), vp(kd_pos), vp(kd_rect),
                                    None if deg is None else vp(deg), self.device,
                                    ctypes.byref(self._m))
        _lib.check(rc, "nat_mesh_from_packed")
        self._finalizer = weakref.finalize(self, L.nat_mesh_free, self._m)
        return self

    def get_weights(self, queryPoints):
        queryPoints = _as_c_double_2d(queryPoints)
        numQueryPoints = queryPoints.shape[0]
        if queryPoints.shape[1] != 3:
            raise ValueError("Query points must have shape (num_query_points, 3)")
        L = _lib.lib()
        rowptr = np.empty(numQueryPoints + 1, dtype=np.int32)
        vals = ctypes.POINTER(ctypes.c_double)()
        cols = ctypes.POINTER(ctypes.c_int32)()
        rc = L.nat_get_weights(queryPoints.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               numQueryPoints, self._m, self.numDataPoints,
                               ctypes.byref(vals), ctypes.byref(cols),
                               rowptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        if rc == -1:
            raise OverflowError("Total number of nonzero weights exceeds int32 limit")
        _lib.check(rc, "nat_get_weights")
        nnz = int(rowptr[numQueryPoints])
        v = _adopt(vals, nnz, ctypes.c_double, np.float64)
        c = _adopt(cols, nnz, ctypes.c_int32, np.int32)
        return scipy.sparse.csr_matrix((v, c, rowptr),
                                       shape=(numQueryPoints, self.numDataPoints))

    # ---- device-resident forms (PyTorch tensors on this mesh's GPU)

    def compute_device(self, queries):
        """The rows of centred float64 CUDA queries [M,3], kept in the mesh's
        row pool for ``weights_spmm`` (no CSR emitted); returns M.  Raises
        OverflowError, as ``get_weights`` would, past INT32_MAX entries."""
        import torch
        q = queries.detach().to(dtype=torch.float64).contiguous()
        if q.dim() != 2 or q.shape[1] != 3:
            raise ValueError("Query points must have shape (num_query_points, 3)")
        stream = torch.cuda.current_stream(q.device).cuda_stream
        nnz = ctypes.c_int64()
        _lib.check(_lib.lib().nat_weights_compute(self._m, ctypes.c_void_p(q.data_ptr()),
                                                  q.shape[0], ctypes.c_void_p(stream),
                                                  ctypes.byref(nnz)), "nat_weights_compute")
        if nnz.value > 2147483647:
            raise OverflowError("Total number of nonzero weights exceeds int32 limit")
        return q.shape[0]

    def weights_device(self, queries):
        """CSR of centred float64 CUDA queries [M,3] as three device tensors
        (crow int32[M+1], col int32[nnz], val float64[nnz])."""
        import torch
        q = queries.detach().to(dtype=torch.float64).contiguous()
        if q.dim() != 2 or q.shape[1] != 3:
            raise ValueError("Query points must have shape (num_query_points, 3)")
        M = q.shape[0]
        L = _lib.lib()
        stream = torch.cuda.current_stream(q.device).cuda_stream
        nnz = ctypes.c_int64()
        rc = L.nat_weights_compute(self._m, ctypes.c_void_p(q.data_ptr()), M,
                                   ctypes.c_void_p(stream), ctypes.byref(nnz))
        _lib.check(rc, "nat_weights_compute")
        n = int(nnz.value)
        if n > 2147483647:
            raise OverflowError("Total number of nonzero weights exceeds int32 limit")
        val = torch.empty(n, dtype=torch.float64, device=q.device)
        col = torch.empty(n, dtype=torch.int32, device=q.device)
        crow = torch.empty(M + 1, dtype=torch.int32, device=q.device)
        rc = L.nat_weights_emit(self._m, ctypes.c_void_p(val.data_ptr() if n else 0),
                                ctypes.c_void_p(col.data_ptr() if n else 0),
                                ctypes.c_void_p(crow.data_ptr()), ctypes.c_void_p(stream))
        if rc == -1:
            raise OverflowError("Total number of nonzero weights exceeds int32 limit")
        _lib.check(rc, "nat_weights_emit")
        return crow, col, val


def row_order(handle, num_rows):
    """Device pointer (ctypes.c_void_p) to the query order of the last weights
    compute on the mesh ``handle`` (neighbouring queries close in space), for
    ``spmm_device``; valid until the next compute on that mesh."""
    out = ctypes.c_void_p()
    _lib.check(_lib.lib().nat_weights_row_order(handle, num_rows, ctypes.byref(out)),
               "nat_weights_row_order")
    return out


def weights_spmm(handle, num_rows, values):
    """``W @ values`` (float64 [M,D]) for the rows of the last weights compute
    on the mesh ``handle``, read from the mesh's row pool: bit for bit
    ``spmm_device`` on the emitted CSR.  values [N,D] float32/float64 CUDA,
    D = 1..4."""
    import torch
    codes = {torch.float32: 0, torch.float64: 1}
    values = values.contiguous()
    out = torch.empty((num_rows, values.shape[1]), dtype=torch.float64, device=values.device)
    stream = torch.cuda.current_stream(values.device).cuda_stream
    _lib.check(_lib.lib().nat_weights_spmm(
        handle, num_rows, ctypes.c_void_p(values.data_ptr()), values.shape[0],
        codes[values.dtype], values.shape[1], ctypes.c_void_p(out.data_ptr()),
        ctypes.c_void_p(stream)), "nat_weights_spmm")
    return out


def spmm_device(crow, col, val, values, num_rows, order=None):
    """``W @ values`` on the device with scipy's accumulation order and dtype
    rule; values [N,D] float32/float64, val float32/float64.  ``order``
    (``row_order``) only sets the order rows are visited in."""
    import torch
    out_dt = torch.float64 if (val.dtype == torch.float64 or values.dtype == torch.float64) \
        else torch.float32
    D = values.shape[1]
    out = torch.empty((num_rows, D), dtype=out_dt, device=values.device)
    codes = {torch.float32: 0, torch.float64: 1}
    stream = torch.cuda.current_stream(values.device).cuda_stream
    rc = _lib.lib().nat_spmm_csr(
        num_rows, values.shape[0], D, ctypes.c_void_p(crow.data_ptr()),
        ctypes.c_void_p(col.data_ptr() if col.numel() else 0),
        ctypes.c_void_p(val.data_ptr() if val.numel() else 0), codes[val.dtype],
        ctypes.c_void_p(values.data_ptr()), codes[values.dtype], order,
        ctypes.c_void_p(out.data_ptr()), codes[out_dt], ctypes.c_void_p(stream))
    _lib.check(rc, "nat_spmm_csr")
    return out


def _as_c_double_2d(a):
    if not isinstance(a, np.ndarray):
        raise TypeError(f"expected a numpy.ndarray, got {type(a).__name__}")
    if a.ndim != 2:
        raise ValueError(f"Buffer has wrong number of dimensions (expected 2, got {a.ndim})")
    if a.dtype != np.float64:
        raise ValueError(f"Buffer dtype mismatch, expected 'double' but got '{a.dtype}'")
    if not a.flags.c_contiguous:
        raise ValueError("ndarray is not C-contiguous")
    return a


def interpolate(
    queries: np.ndarray,
    keys: np.ndarray,
    values: np.ndarray,
    parallel: bool = True,
    num_threads: Optional[int] = None,
) -> np.ndarray:
    """Natural neighbour (Sibson) interpolation of ``values`` known at ``keys``
    [N,3], evaluated at ``queries`` [M,3]; values [N] or [N,D] -> [M] or [M,D].
    Same contract as natinterp3d.interpolate (natinterp3d.py:9-35); the mesh of
    recently used keys is reused (see ``clear_mesh_cache``)."""
    interpolator = _cached_interpolator(keys, parallel, num_threads)
    return interpolator.interpolate(queries, values)


def get_weights(
    queries: np.ndarray,
    keys: np.ndarray,
    parallel: bool = True,
    num_threads: Optional[int] = None,
) -> scipy.sparse.csr_matrix:
    """Sibson coordinates of ``queries`` with respect to ``keys`` as a CSR
    matrix [M,N].  Same contract as natinterp3d.get_weights
    (natinterp3d.py:38-60); the mesh of recently used keys is reused."""
    interpolator = _cached_interpolator(keys, parallel, num_threads)
    return interpolator.get_weights(queries)


# Device meshes of the keys of recent module-level calls, most recent last.
# The reference builds a mesh per call (natinterp3d.py:34, 59); the build is
# deterministic, so a mesh built earlier for the same keys gives the same
# rows.  Size from $NATINTERP3D_MESH_CACHE (default 2; 0 turns it off).
_MESH_CACHE = collections.OrderedDict()
_MESH_CACHE_STATS = {"hits": 0, "misses": 0}


def _cached_interpolator(keys, parallel, num_threads):
    size = int(os.environ.get("NATINTERP3D_MESH_CACHE", "2"))
    if size <= 0:
        return Interpolator(keys, parallel=parallel, num_threads=num_threads)
    k = np.ascontiguousarray(keys, np.float64)
    dev = _default_device()
    key = (k.shape, hashlib.blake2b(memoryview(k).cast("B"), digest_size=20).hexdigest(), dev)
    it = _MESH_CACHE.get(key)
    if it is not None:
        _MESH_CACHE.move_to_end(key)
        _MESH_CACHE_STATS["hits"] += 1
        return it
    _MESH_CACHE_STATS["misses"] += 1
    it = Interpolator(k, parallel=parallel, num_threads=num_threads, device=dev)
    _MESH_CACHE[key] = it
    while len(_MESH_CACHE) > size:
        _MESH_CACHE.popitem(last=False)
    return it


def clear_mesh_cache():
    """Drop the device meshes kept for module-level calls."""
    _MESH_CACHE.clear()


def mesh_cache_stats():
    """Hits and misses of the module-level mesh cache, and its entries."""
    return {**_MESH_CACHE_STATS, "entries": len(_MESH_CACHE)}


class Interpolator:
    """Natural neighbour interpolator for 3D data points (natinterp3d.py:63-137).

    The Delaunay mesh of ``data_points`` is built once and kept in GPU memory;
    reuse the object for repeated queries on the same keys.  ``device``
    selects the GPU (default: the current PyTorch device, or
    ``$NATINTERP3D_DEVICE``)."""

    def __init__(
        self,
        data_points: np.ndarray,
        parallel: bool = True,
        num_threads: Optional[int] = None,
        device: Optional[int] = None,
    ):
        data_points = np.ascontiguousarray(data_points, np.float64)
        self._centroid = data_points.mean(axis=0)
        centered = data_points - self._centroid
        self.parallel = parallel
        self.num_threads = num_threads
        self.mesh_and_vertices = MeshAndVerticesHIP(centered, device=device)

    def save(self, path):
        """Write the mesh and the key centroid to ``path`` (.npz, plain arrays):
        ``Interpolator.load`` then skips the host build."""
        self.mesh_and_vertices.save(path, centroid=np.asarray(self._centroid, np.float64))

    @classmethod
    def load(cls, path, parallel: bool = True, num_threads: Optional[int] = None,
             device: Optional[int] = None) -> "Interpolator":
        """The Interpolator saved by ``save``; its rows equal the original's."""
        with np.load(path, allow_pickle=False) as z:
            if "centroid" not in z:
                raise ValueError(f"{path}: not a saved Interpolator (no centroid)")
            centroid = np.asarray(z["centroid"], np.float64).reshape(3)
        self = cls.__new__(cls)
        self._centroid = centroid
        self.parallel = parallel
        self.num_threads = num_threads
        self.mesh_and_vertices = MeshAndVerticesHIP.load(path, device=device)
        return self

    def autotune(self, query_points: np.ndarray, repeats: int = 3) -> dict:
        """Time the row-path policies on representative ``query_points`` and
        keep the fastest for later calls (results never depend on it)."""
        import torch
        q = np.ascontiguousarray(query_points, np.float64) - self._centroid
        dev = torch.device("cuda", self.mesh_and_vertices.device)
        with torch.cuda.device(dev):
            return self.mesh_and_vertices.autotune(torch.from_numpy(q).to(dev), repeats)

    def get_weights(self, query_points: np.ndarray) -> scipy.sparse.csr_matrix:
        """Sibson coordinates of ``query_points`` as a CSR matrix whose data has
        the dtype of ``query_points`` (natinterp3d.py:105-117)."""
        out_dtype = query_points.dtype
        query_points = np.ascontiguousarray(query_points, np.float64) - self._centroid
        result = self.mesh_and_vertices.get_weights(query_points)
        return result.astype(out_dtype, copy=False)

    def interpolate(self, query_points: np.ndarray, values: np.ndarray) -> np.ndarray:
        """``get_weights(query_points) @ values`` (natinterp3d.py:119-137).  The
        weights never leave the GPU: the CSR is built and multiplied there, in
        scipy's accumulation order and result dtype."""
        import torch
        out_dtype = np.dtype(query_points.dtype)
        values = np.asarray(values)
        one_d = len(values.shape) == 1
        vals2 = values[:, np.newaxis] if one_d else values
        if vals2.ndim != 2:
            raise ValueError("values must have shape (N,) or (N, D)")
        n = self.mesh_and_vertices.numDataPoints
        if vals2.shape[0] != n:
            raise ValueError("dimension mismatch")
        res_dtype = np.result_type(out_dtype, vals2.dtype)
        if out_dtype not in _DT_CODE or res_dtype not in _DT_CODE:
            raise TypeError(
                f"GPU interpolation supports float32/float64 weights and results "
                f"(query dtype {out_dtype}, values dtype {vals2.dtype})")
        q = np.ascontiguousarray(query_points, np.float64) - self._centroid
        if q.ndim != 2 or q.shape[1] != 3:
            raise ValueError("Query points must have shape (num_query_points, 3)")
        dev = torch.device("cuda", self.mesh_and_vertices.device)
        with torch.cuda.device(dev):
            qd = torch.from_numpy(q).to(dev)
            vdt = vals2.dtype if vals2.dtype in _DT_CODE else res_dtype
            if out_dtype == np.float64 and 1 <= vals2.shape[1] <= 4:
                # float64 weights: the product straight from the row pool,
                # without emitting the CSR (same sums, same order)
                M = self.mesh_and_vertices.compute_device(qd)
                xd = torch.from_numpy(np.ascontiguousarray(vals2, vdt)).to(dev)
                out = weights_spmm(self.mesh_and_vertices.handle, M, xd).cpu().numpy()
                if out.dtype != res_dtype:
                    out = out.astype(res_dtype)
                return np.squeeze(out, axis=1) if one_d else out
            crow, col, val = self.mesh_and_vertices.weights_device(qd)
            if out_dtype != np.float64:
                # get_weights' astype(out_dtype) canonicalises W (rows sorted
                # by column) before the product; do the same on the device
                val = val.to(torch.float32)
                stream = torch.cuda.current_stream(dev).cuda_stream
                _lib.check(_lib.lib().nat_csr_sort_rows(
                    q.shape[0], ctypes.c_void_p(crow.data_ptr()),
                    ctypes.c_void_p(col.data_ptr() if col.numel() else 0),
                    ctypes.c_void_p(val.data_ptr() if val.numel() else 0), 0,
                    ctypes.c_void_p(stream)), "nat_csr_sort_rows")
            xd = torch.from_numpy(np.ascontiguousarray(vals2, vdt)).to(dev)
            y = spmm_device(crow, col, val, xd, q.shape[0],
                            order=row_order(self.mesh_and_vertices.handle, q.shape[0]))
            out = y.cpu().numpy()
        if out.dtype != res_dtype:
            out = out.astype(res_dtype)
        return np.squeeze(out, axis=1) if one_d else out

    def interpolate_torch(self, query_points, values):
        """Differentiable ``W(query_points) @ values`` for torch tensors
        (query_points [M,3], values [N] or [N,D], floating point).  The
        weights are the float64 Sibson coordinates of ``get_weights``; the
        product accumulates in float64 and is returned in the promoted dtype
        of the two inputs.  Gradients flow to both inputs (see
        natinterp3d_amd.autograd)."""
        import torch

        from .autograd import SibsonInterpolation
        dev = torch.device("cuda", self.mesh_and_vertices.device)
        if not (query_points.is_floating_point() and values.is_floating_point()):
            raise TypeError("query_points and values must be floating point tensors")
        one_d = values.dim() == 1
        v2 = values.unsqueeze(1) if one_d else values
        if v2.dim() != 2 or v2.shape[0] != self.mesh_and_vertices.numDataPoints:
            raise ValueError("values must have shape (N,) or (N, D)")
        if query_points.dim() != 2 or query_points.shape[1] != 3:
            raise ValueError("Query points must have shape (num_query_points, 3)")
        centroid = torch.as_tensor(self._centroid, dtype=torch.float64, device=dev)
        q = query_points.to(device=dev, dtype=torch.float64) - centroid
        y = SibsonInterpolation.apply(self.mesh_and_vertices, q, v2.to(dev))
        y = y.to(torch.promote_types(query_points.dtype, values.dtype))
        return y.squeeze(1) if one_d else y


class WeightsGraph:
    """A weights computation captured once in a hipGraph and replayed.

    ``WeightsGraph(mesh, queries[, query_offsets])`` runs one computation of
    the centred float64 CUDA ``queries`` [M,3] (a batch mesh also takes its
    query offsets) to size the workspace, then captures the launch sequence.
    ``run()`` replays it for whatever ``queries`` holds at that moment (update
    it in place) and returns the CSR (crow, col, val) as new tensors, or
    raises ``GraphCapacityError`` when the new queries need more workspace
    than the graph was made with (then compute them with ``weights_device``
    or make a new graph), and ``GraphPrecisionError`` when the mesh's
    precision was changed after the capture."""

    def __init__(self, mesh, queries, query_offsets=None):
        import torch
        if queries.dtype != torch.float64 or not queries.is_cuda or not queries.is_contiguous():
            raise ValueError("queries must be a contiguous float64 CUDA tensor")
        if queries.dim() != 2 or queries.shape[1] != 3 or queries.shape[0] == 0:
            raise ValueError("queries must have shape (num_query_points, 3), M > 0")
        self._mesh = mesh
        self.queries = queries
        self.M = queries.shape[0]
        L = _lib.lib()
        off = None
        if query_offsets is not None:
            self._off = np.ascontiguousarray(query_offsets, np.int64)
            off = self._off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        stream = ctypes.c_void_p(torch.cuda.current_stream(queries.device).cuda_stream)
        self._g = ctypes.c_void_p()
        _lib.check(L.nat_graph_create(mesh.handle, ctypes.c_void_p(queries.data_ptr()), self.M,
                                      off, stream, ctypes.byref(self._g)), "nat_graph_create")
        self._finalizer = weakref.finalize(self, L.nat_graph_free, self._g)

    def run(self):
        import torch
        dev = self.queries.device
        L = _lib.lib()
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        nnz = ctypes.c_int64()
        rc = L.nat_graph_launch(self._g, stream, ctypes.byref(nnz))
        if rc == -8:
            raise GraphCapacityError(_lib.last_error())
        if rc == -11:
            raise GraphPrecisionError(_lib.last_error())
        _lib.check(rc, "nat_graph_launch")
        n = int(nnz.value)
        val = torch.empty(n, dtype=torch.float64, device=dev)
        col = torch.empty(n, dtype=torch.int32, device=dev)
        crow = torch.empty(self.M + 1, dtype=torch.int32, device=dev)
        _lib.check(L.nat_graph_copy_out(self._g, ctypes.c_void_p(val.data_ptr() if n else 0),
                                        ctypes.c_void_p(col.data_ptr() if n else 0),
                                        ctypes.c_void_p(crow.data_ptr()), stream),
                   "nat_graph_copy_out")
        return crow, col, val


class GraphCapacityError(RuntimeError):
    """A graph replay needed more workspace than the graph was made with."""


class GraphPrecisionError(RuntimeError):
    """A graph replay after ``set_precision`` changed the mesh's row
    arithmetic: the graph holds the kernels of its capture (make a new one)."""


def get_row_policy(handle, num_problems):
    """Row-path policy of a mesh handle as a dict (budgets per problem)."""
    pol = _lib.Policy()
    budgets = np.zeros(num_problems, np.uint8)
    _lib.check(_lib.lib().nat_mesh_get_policy(handle, ctypes.byref(pol),
                                              budgets.ctypes.data_as(ctypes.c_void_p)),
               "nat_mesh_get_policy")
    return {"budgets": budgets.tolist(), "first_level": pol.first_level,
            "work_sort": bool(pol.work_sort), "topo_grid": pol.topo_grid,
            "geom_grid": pol.geom_grid, "wave_rows": bool(pol.wave_rows),
            "geom_lanes": pol.geom_lanes}


def set_row_policy(handle, num_problems, budgets=None, first_level=None, work_sort=None,
                   topo_grid=None, geom_grid=None, wave_rows=None, geom_lanes=None):
    """Set the given fields of a mesh handle's row-path policy."""
    cur = get_row_policy(handle, num_problems)
    pol = _lib.Policy()
    pol.first_level = int(cur["first_level"] if first_level is None else first_level)
    pol.work_sort = int(bool(cur["work_sort"] if work_sort is None else work_sort))
    pol.topo_grid = int(cur["topo_grid"] if topo_grid is None else topo_grid)
    pol.geom_grid = int(cur["geom_grid"] if geom_grid is None else geom_grid)
    pol.wave_rows = int(bool(cur["wave_rows"] if wave_rows is None else wave_rows))
    pol.geom_lanes = int(cur["geom_lanes"] if geom_lanes is None else geom_lanes)
    b = None
    if budgets is not None:
        arr = np.asarray(budgets)
        if arr.shape != (num_problems,) or (arr < 0).any() or (arr > 255).any():
            raise ValueError(f"budgets must be {num_problems} values in 0..96")
        b = np.ascontiguousarray(arr, np.uint8)
    _lib.check(_lib.lib().nat_mesh_set_policy(
        handle, ctypes.byref(pol), None if b is None else b.ctypes.data_as(ctypes.c_void_p)),
        "nat_mesh_set_policy")


def autotune_rows(handle, queries, query_offsets, repeats):
    """nat_mesh_autotune on centred float64 CUDA queries (query_offsets for a
    batch mesh); returns the report as a dict."""
    import torch
    q = queries.detach().to(dtype=torch.float64).contiguous()
    off = None
    if query_offsets is not None:
        off = np.ascontiguousarray(query_offsets, np.int64)
    rep = _lib.TuneReport()
    stream = ctypes.c_void_p(torch.cuda.current_stream(q.device).cuda_stream)
    _lib.check(_lib.lib().nat_mesh_autotune(
        handle, ctypes.c_void_p(q.data_ptr() if q.numel() else 0), q.shape[0],
        None if off is None else off.ctypes.data_as(ctypes.c_void_p), int(repeats), stream,
        ctypes.byref(rep)), "nat_mesh_autotune")
    return {f: getattr(rep, f) for f, _ in rep._fields_}
