"""Differentiable Sibson interpolation on the GPU (no reference counterpart:
natinterp3d has no gradients).

``SibsonInterpolation.apply(mesh, queries, values)`` computes
``y = W(queries) @ values`` for centred float64 CUDA queries [M,3] and values
[N,D]; its backward gives

* dL/dvalues = W^T dL/dy  (float64 atomics, nat_spmm_t_csr), and
* dL/dqueries from the analytic derivative of the Sibson coordinates: with
  sigma_j the volume query i steals from neighbour j, S = sum_j sigma_j and
  lambda_j = sigma_j / S,  d sigma_j / dq = (moment about q of the Voronoi
  facet shared by q and p_j) / |p_j - q|  and
  d lambda_j / dq = (d sigma_j/dq - lambda_j sum_k d sigma_k/dq) / S.

The forward pass keeps each row's cavity, boundary faces and neighbour order
(nat_topology_save); the backward pass recomputes the circumcentres and
volumes from them rather than storing per-entry derivatives.  Empty rows,
rows of queries that coincide with a data point and rows outside the hull's
super simplex have zero query gradient.
"""
import ctypes

import torch

from . import _lib
from .natinterp3d import row_order, spmm_device, weights_spmm


# saved-topology bytes per query of the last forward on each mesh (by handle)
_TOPO_BYTES_PER_QUERY = {}


_SIDE = {}


def _side_stream(device):
    """One extra stream per device for the backward's W^T g."""
    key = torch.device(device).index
    if key not in _SIDE:
        _SIDE[key] = torch.cuda.Stream(device=device)
    return _SIDE[key]


def _vp(t):
    return ctypes.c_void_p(t.data_ptr() if t is not None and t.numel() else 0)


class SibsonInterpolation(torch.autograd.Function):
    """y = W(queries) @ values with gradients to queries and values."""

    @staticmethod
    def forward(ctx, mesh, queries, values):
        if not queries.is_cuda or not values.is_cuda:
            raise ValueError("queries and values must be CUDA tensors")
        if queries.dim() != 2 or queries.shape[1] != 3:
            raise ValueError("Query points must have shape (num_query_points, 3)")
        if values.dim() != 2 or values.shape[0] != mesh.numDataPoints:
            raise ValueError("values must have shape (num_data_points, D)")
        q = queries.detach().to(torch.float64).contiguous()
        v = values.detach().to(torch.float64).contiguous()
        M = q.shape[0]
        L = _lib.lib()
        topo = None
        key = mesh.handle.value
        if ctx.needs_input_grad[1] and M:
            # the compute writes its records straight into the saved topology
            # (sized from the last one of this mesh, with a margin); if they
            # do not fit, the save below copies them into a larger buffer
            est = int(_TOPO_BYTES_PER_QUERY.get(key, 700.0) * 1.15 * M) + 4096
            topo = torch.empty(est, dtype=torch.uint8, device=q.device)
            _lib.check(L.nat_topology_target(mesh.handle, ctypes.c_void_p(topo.data_ptr()),
                                             topo.numel(), M), "nat_topology_target")
        try:
            crow, col, val = mesh.weights_device(q)
        finally:
            if topo is not None:  # consumed by the compute; cleared if it never ran
                L.nat_topology_target(mesh.handle, None, 0, 0)
        if ctx.needs_input_grad[1]:
            nbytes = ctypes.c_int64()
            _lib.check(L.nat_topology_save(mesh.handle, None, 0, ctypes.byref(nbytes), None),
                       "nat_topology_save")
            if topo is None or int(nbytes.value) > topo.numel():
                topo = torch.empty(int(nbytes.value), dtype=torch.uint8, device=q.device)
            stream = torch.cuda.current_stream(q.device).cuda_stream
            _lib.check(L.nat_topology_save(mesh.handle, ctypes.c_void_p(topo.data_ptr()),
                                           topo.numel(), ctypes.byref(nbytes),
                                           ctypes.c_void_p(stream)), "nat_topology_save")
            if M:
                _TOPO_BYTES_PER_QUERY[key] = int(nbytes.value) / M
        if 1 <= v.shape[1] <= 4:  # from the row pool: same sums, read in order
            y = weights_spmm(mesh.handle, M, v)
        else:
            y = spmm_device(crow, col, val, v, M, order=row_order(mesh.handle, M))
        ctx.mesh = mesh
        ctx.values_dtype = values.dtype
        ctx.queries_dtype = queries.dtype
        ctx.save_for_backward(q, v, crow, col, val, topo)
        return y

    @staticmethod
    def backward(ctx, gy):
        q, v, crow, col, val, topo = ctx.saved_tensors
        mesh = ctx.mesh
        gy = gy.to(torch.float64).contiguous()
        M, D = gy.shape
        L = _lib.lib()
        cur = torch.cuda.current_stream(gy.device)
        stream = ctypes.c_void_p(cur.cuda_stream)
        dq = dv = None
        if ctx.needs_input_grad[2]:
            # W^T g on a second stream, beside the query gradient: that
            # kernel's resident blocks leave most of each CU's LDS and wave
            # slots free (default priority, like every stream here)
            order = ctypes.c_void_p()
            if topo is not None and M:
                _lib.check(L.nat_topology_row_order(ctypes.c_void_p(topo.data_ptr()), M,
                                                    ctypes.byref(order)),
                           "nat_topology_row_order")
            side = _side_stream(gy.device)
            side.wait_stream(cur)
            with torch.cuda.stream(side):
                dv = torch.empty(v.shape, dtype=torch.float64, device=gy.device)
                _lib.check(L.nat_spmm_t_csr(M, v.shape[0], D, _vp(crow), _vp(col), _vp(val),
                                            _vp(gy), order, _vp(dv),
                                            ctypes.c_void_p(side.cuda_stream)),
                           "nat_spmm_t_csr")
            for t in (crow, col, val, gy):
                t.record_stream(side)
        if ctx.needs_input_grad[1]:
            dq = torch.empty((M, 3), dtype=torch.float64, device=gy.device)
            if M:
                _lib.check(L.nat_grad_queries(mesh.handle, ctypes.c_void_p(topo.data_ptr()),
                                              _vp(q), M, _vp(crow), _vp(val), _vp(gy), _vp(v),
                                              D, _vp(dq), stream), "nat_grad_queries")
            dq = dq.to(ctx.queries_dtype)
        if dv is not None:
            cur.wait_stream(side)
            dv.record_stream(cur)
            dv = dv.to(ctx.values_dtype)
        return None, dq, dv
