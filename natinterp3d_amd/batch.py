"""Many independent interpolation problems in the same GPU launches (no
reference counterpart: natinterp3d handles one key set per Interpolator).

``BatchInterpolator(key_sets)`` builds every problem's mesh on the host (in
parallel threads) and uploads them as one mesh with global vertex numbering.
Each call then runs the queries of all problems through one pass of the
kernels -- one launch per kernel for the whole batch -- and splits the result
per problem.  Problem p's rows, weights and products are exactly those of
``Interpolator(key_sets[p])`` on the same queries.
"""
import ctypes
import weakref
from typing import List, Optional, Sequence

import numpy as np
import scipy.sparse

from . import _lib
from .natinterp3d import (_DT_CODE, _default_device, autotune_rows, get_row_policy,
                          row_order, set_row_policy, spmm_device)


class BatchInterpolator:
    """Independent natural-neighbour problems sharing one device mesh."""

    def __init__(self, key_sets: Sequence[np.ndarray], device: Optional[int] = None,
                 num_threads: Optional[int] = None):
        keys = [np.ascontiguousarray(k, np.float64) for k in key_sets]
        if not keys:
            raise ValueError("need at least one problem")
        for p, k in enumerate(keys):
            if k.ndim != 2 or k.shape[1] != 3 or k.shape[0] == 0:
                raise ValueError(f"problem {p}: keys must have shape (num_keys, 3)")
        self._centroids = [k.mean(axis=0) for k in keys]
        self._centered = [np.ascontiguousarray(k - c) for k, c in zip(keys, self._centroids)]
        self.num_problems = len(keys)
        self.num_keys = [k.shape[0] for k in keys]
        self.device = _default_device() if device is None else int(device)
        L = _lib.lib()
        dp = ctypes.POINTER(ctypes.c_double)
        ptrs = (dp * self.num_problems)(*[c.ctypes.data_as(dp) for c in self._centered])
        counts = (ctypes.c_int32 * self.num_problems)(*self.num_keys)
        self._m = ctypes.c_void_p()
        rc = L.nat_batch_build(ptrs, counts, self.num_problems, self.device,
                               int(num_threads or 0), ctypes.byref(self._m))
        _lib.check(rc, "nat_batch_build")
        self._finalizer = weakref.finalize(self, L.nat_mesh_free, self._m)
        voff = (ctypes.c_int64 * (self.num_problems + 1))()
        _lib.check(L.nat_batch_offsets(self._m, voff, None), "nat_batch_offsets")
        self.vertex_offsets = np.array(voff[:], dtype=np.int64)

    @property
    def handle(self):
        return self._m

    def info(self):
        inf = _lib.MeshInfo()
        _lib.check(_lib.lib().nat_mesh_get_info(self._m, ctypes.byref(inf)), "nat_mesh_get_info")
        return {f: getattr(inf, f) for f, _ in inf._fields_}

    def set_timing(self, enable=True):
        _lib.check(_lib.lib().nat_mesh_set_timing(self._m, 1 if enable else 0),
                   "nat_mesh_set_timing")

    def policy(self):
        """The row-path policy, budgets per problem (see set_policy)."""
        return get_row_policy(self._m, self.num_problems)

    def set_policy(self, **kw):
        """Change the row-path policy: budgets (per problem, 0..96; 0 = skip
        the group kernel), first_level (1 or 2), work_sort (bool), topo_grid,
        geom_grid (1..4), wave_rows (bool), geom_lanes (2 or 1).  Results do
        not depend on it."""
        set_row_policy(self._m, self.num_problems, **kw)

    def autotune(self, query_sets, repeats=3):
        """Time row-path policies on these per-problem query sets and keep the
        fastest (problems whose queries mostly overflow the group kernel's
        LDS tables go straight to thread-per-query rows, and so on); returns
        the tuner's report (dict)."""
        import torch
        q, off = self._centre(query_sets)
        dev = torch.device("cuda", self.device)
        with torch.cuda.device(dev):
            return autotune_rows(self._m, torch.from_numpy(q).to(dev), off, repeats)

    def autotune_device(self, queries, query_offsets, repeats=3):
        """autotune on centred float64 CUDA queries with their offsets."""
        return autotune_rows(self._m, queries, query_offsets, repeats)

    def _centre(self, query_sets):
        if len(query_sets) != self.num_problems:
            raise ValueError(f"expected {self.num_problems} query sets, got {len(query_sets)}")
        parts = []
        for p, q in enumerate(query_sets):
            q = np.asarray(q)
            if q.ndim != 2 or q.shape[1] != 3:
                raise ValueError(f"problem {p}: queries must have shape (num_queries, 3)")
            parts.append(np.ascontiguousarray(q, np.float64) - self._centroids[p])
        offsets = np.zeros(self.num_problems + 1, np.int64)
        offsets[1:] = np.cumsum([len(q) for q in parts])
        return np.ascontiguousarray(np.concatenate(parts, axis=0)), offsets

    def weights_device(self, queries, query_offsets):
        """Block-diagonal CSR of centred float64 CUDA queries (all problems
        back to back, problem p's at rows query_offsets[p]:query_offsets[p+1])
        as device tensors (crow int32, col int32 global, val float64)."""
        import torch
        q = queries.detach().to(dtype=torch.float64).contiguous()
        off = np.ascontiguousarray(query_offsets, np.int64)
        if off.shape != (self.num_problems + 1,) or off[-1] != q.shape[0]:
            raise ValueError("query_offsets must have num_problems + 1 entries ending at M")
        L = _lib.lib()
        stream = torch.cuda.current_stream(q.device).cuda_stream
        nnz = ctypes.c_int64()
        rc = L.nat_batch_weights_compute(
            self._m, ctypes.c_void_p(q.data_ptr() if q.numel() else 0),
            off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.c_void_p(stream),
            ctypes.byref(nnz))
        _lib.check(rc, "nat_batch_weights_compute")
        n = int(nnz.value)
        if n > 2147483647:
            raise OverflowError("Total number of nonzero weights exceeds int32 limit")
        M = q.shape[0]
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

    def interpolate_torch(self, queries, query_offsets, values):
        """Differentiable batched ``W @ values`` on torch tensors: queries [M,3]
        (all problems back to back, uncentred), query_offsets [P+1], values
        [sum N_p, D] (problem p's keys at vertex_offsets[p]).  Returns [M, D]
        in the promoted dtype; gradients flow to queries and values."""
        import torch

        from .autograd import SibsonInterpolation
        dev = torch.device("cuda", self.device)
        off = np.ascontiguousarray(query_offsets, np.int64)
        counts = np.diff(off)
        cent = torch.as_tensor(np.repeat(np.asarray(self._centroids).reshape(-1, 3), counts,
                                         axis=0), dtype=torch.float64, device=dev)
        q = queries.to(device=dev, dtype=torch.float64) - cent
        y = SibsonInterpolation.apply(_BatchCall(self, off), q, values.to(dev))
        return y.to(torch.promote_types(queries.dtype, values.dtype))

    def get_weights(self, query_sets) -> List[scipy.sparse.csr_matrix]:
        """Per problem, the CSR matrix Interpolator(keys_p).get_weights(queries_p)
        returns: float64 Sibson coordinates cast to the queries' dtype, local
        column indices in natural-neighbour discovery order."""
        import torch
        dtypes = [np.asarray(q).dtype for q in query_sets]
        q, off = self._centre(query_sets)
        dev = torch.device("cuda", self.device)
        with torch.cuda.device(dev):
            crow, col, val = self.weights_device(torch.from_numpy(q).to(dev), off)
            crow, col, val = crow.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy()
        out = []
        for p in range(self.num_problems):
            r0, r1 = off[p], off[p + 1]
            e0, e1 = crow[r0], crow[r1]
            W = scipy.sparse.csr_matrix(
                (val[e0:e1], (col[e0:e1] - self.vertex_offsets[p]).astype(np.int32),
                 (crow[r0:r1 + 1] - e0).astype(np.int32)),
                shape=(int(r1 - r0), self.num_keys[p]))
            out.append(W.astype(dtypes[p], copy=False))
        return out

    def interpolate(self, query_sets, value_sets) -> List[np.ndarray]:
        """Per problem, Interpolator(keys_p).interpolate(queries_p, values_p),
        with the products of all problems done in one launch.  Every problem
        must use the same query dtype, value dtype and value shape."""
        import torch
        if len(value_sets) != self.num_problems:
            raise ValueError(f"expected {self.num_problems} value sets")
        qdt = {np.asarray(q).dtype for q in query_sets}
        vals = [np.asarray(v) for v in value_sets]
        vdt = {v.dtype for v in vals}
        trailing = {v.shape[1:] for v in vals}
        if len(qdt) != 1 or len(vdt) != 1 or len(trailing) != 1:
            raise ValueError("all problems must share query dtype, value dtype and value shape")
        out_dtype, vdtype = qdt.pop(), vdt.pop()
        one_d = len(trailing.pop()) == 0
        for p, v in enumerate(vals):
            if v.shape[0] != self.num_keys[p]:
                raise ValueError(f"problem {p}: dimension mismatch")
        v2 = [v[:, None] if one_d else v for v in vals]
        if v2[0].ndim != 2:
            raise ValueError("values must have shape (N,) or (N, D)")
        res_dtype = np.result_type(out_dtype, vdtype)
        if out_dtype not in _DT_CODE or res_dtype not in _DT_CODE:
            raise TypeError("GPU interpolation supports float32/float64 weights and results")
        q, off = self._centre(query_sets)
        dev = torch.device("cuda", self.device)
        with torch.cuda.device(dev):
            crow, col, val = self.weights_device(torch.from_numpy(q).to(dev), off)
            if out_dtype != np.float64:
                # per-problem astype canonicalisation (rows sorted by column);
                # global columns sort like local ones within a row
                val = val.to(torch.float32)
                stream = torch.cuda.current_stream(dev).cuda_stream
                _lib.check(_lib.lib().nat_csr_sort_rows(
                    q.shape[0], ctypes.c_void_p(crow.data_ptr()),
                    ctypes.c_void_p(col.data_ptr() if col.numel() else 0),
                    ctypes.c_void_p(val.data_ptr() if val.numel() else 0), 0,
                    ctypes.c_void_p(stream)), "nat_csr_sort_rows")
            vd = vdtype if vdtype in _DT_CODE else res_dtype
            xd = torch.from_numpy(np.ascontiguousarray(np.concatenate(v2, axis=0), vd)).to(dev)
            y = spmm_device(crow, col, val, xd, q.shape[0],
                            order=row_order(self.handle, q.shape[0])).cpu().numpy()
        if y.dtype != res_dtype:
            y = y.astype(res_dtype)
        outs = [y[off[p]:off[p + 1]] for p in range(self.num_problems)]
        return [np.squeeze(o, axis=1) if one_d else o for o in outs]


class _BatchCall:
    """A batch bound to one query split, with the interface the autograd
    function uses (handle, numDataPoints, weights_device)."""

    def __init__(self, batch, offsets):
        self._batch = batch
        self._offsets = offsets
        self.handle = batch.handle
        self.numDataPoints = int(batch.vertex_offsets[-1])

    def weights_device(self, queries):
        return self._batch.weights_device(queries, self._offsets)
