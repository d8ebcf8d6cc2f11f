"""ctypes binding of the C ABI declared in include/natinterp3d_hip.h.

The library is required: there is no CPU path.  Importing the package works
without it (so the API can be inspected), but every call that computes raises
``ImportError`` until ``libnatinterp3d_hip.so`` has been built.
"""
import ctypes
import os

from ._build import LIB_PATH

_lib = None

EXPORTS = (
    "nat_mesh_build", "nat_mesh_free", "nat_mesh_get_info", "nat_mesh_export",
    "nat_get_weights", "nat_weights_compute", "nat_weights_emit", "nat_spmm_csr",
    "nat_csr_sort_rows", "nat_topology_save", "nat_topology_row_order", "nat_grad_queries",
    "nat_spmm_t_csr",
    "nat_mesh_set_timing", "nat_mesh_get_timings", "nat_free_host", "nat_last_error",
    "nat_batch_build", "nat_batch_offsets", "nat_batch_weights_compute",
    "nat_mesh_set_precision", "nat_predicate_audit", "nat_insphere_classify",
    "nat_graph_create", "nat_graph_launch", "nat_graph_copy_out", "nat_graph_free",
    "nat_mesh_set_policy", "nat_mesh_get_policy", "nat_mesh_autotune",
    "nat_mesh_from_packed", "nat_weights_row_order", "nat_topology_target",
    "nat_weights_spmm", "nat_weights_noalloc", "nat_mesh_set_predicates",
    "nat_predicate_exact", "nat_mesh_set_mixed_check", "nat_mesh_build_ex",
    "nat_mesh_delaunay_audit",
)


class Timings(ctypes.Structure):
    _fields_ = [
        ("order_ms", ctypes.c_double), ("locate_ms", ctypes.c_double),
        ("topo_ms", ctypes.c_double),
        ("geom_ms", ctypes.c_double), ("rows_ms", ctypes.c_double * 4),
        ("rows_items", ctypes.c_int32 * 4), ("levels", ctypes.c_int32),
        ("emit_ms", ctypes.c_double), ("grad_ms", ctypes.c_double * 4),
        ("grad_items", ctypes.c_int32 * 4), ("mixed_fp32", ctypes.c_int64),
        ("mixed_fp64", ctypes.c_int64), ("mixed_rows_fp64", ctypes.c_int64),
        ("wave_ms", ctypes.c_double), ("grad_wave_ms", ctypes.c_double),
        ("wave_items", ctypes.c_int32), ("grad_wave_items", ctypes.c_int32),
        ("topo_search_ms", ctypes.c_double), ("topo_order_ms", ctypes.c_double),
        ("robust_exact", ctypes.c_int64), ("robust_changed", ctypes.c_int64),
        ("robust_unresolved", ctypes.c_int64), ("mixed_err_max", ctypes.c_double),
        ("mixed_err_rows", ctypes.c_int64), ("mixed_err_entries", ctypes.c_int64),
    ]


class Policy(ctypes.Structure):
    _fields_ = [
        ("first_level", ctypes.c_int32), ("work_sort", ctypes.c_int32),
        ("topo_grid", ctypes.c_int32), ("geom_grid", ctypes.c_int32),
        ("wave_rows", ctypes.c_int32), ("geom_lanes", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 2),
    ]


class TuneReport(ctypes.Structure):
    _fields_ = [
        ("default_ms", ctypes.c_double), ("tuned_ms", ctypes.c_double),
        ("candidates", ctypes.c_int32), ("routed_problems", ctypes.c_int32),
        ("routed_queries", ctypes.c_int64), ("left_group", ctypes.c_double),
    ]


class MeshInfo(ctypes.Structure):
    _fields_ = [
        ("n_data", ctypes.c_int32), ("n_tets", ctypes.c_int32), ("kd_depth", ctypes.c_int32),
        ("coplanar", ctypes.c_int32), ("cospherical", ctypes.c_int32),
        ("edge_anomalies", ctypes.c_int32), ("device", ctypes.c_int32),
        ("num_cus", ctypes.c_int32), ("device_bytes", ctypes.c_int64),
        ("workspace_bytes", ctypes.c_int64), ("n_problems", ctypes.c_int32),
        ("builder", ctypes.c_int32), ("build_rounds", ctypes.c_int32),
        ("build_seconds", ctypes.c_double),
    ]


def lib():
    """The loaded HIP library; raises ImportError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"natinterp3d_amd: {LIB_PATH} is missing; build it with "
            "`python -m natinterp3d_amd._build` (there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    L.nat_mesh_build.restype = ctypes.c_int
    L.nat_mesh_build.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    L.nat_mesh_free.restype = None
    L.nat_mesh_free.argtypes = [vp]
    L.nat_mesh_get_info.restype = ctypes.c_int
    L.nat_mesh_get_info.argtypes = [vp, ctypes.POINTER(MeshInfo)]
    L.nat_mesh_export.restype = ctypes.c_int
    L.nat_mesh_export.argtypes = [vp, dp, ip, ip, ip, ip, dp, dp]
    L.nat_get_weights.restype = ctypes.c_int
    L.nat_get_weights.argtypes = [dp, ctypes.c_int, vp, ctypes.c_int, ctypes.POINTER(dp),
                                  ctypes.POINTER(ip), ip]
    L.nat_weights_compute.restype = ctypes.c_int
    L.nat_weights_compute.argtypes = [vp, vp, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_int64)]
    L.nat_weights_emit.restype = ctypes.c_int
    L.nat_weights_emit.argtypes = [vp, vp, vp, vp, vp]
    L.nat_weights_noalloc.restype = ctypes.c_int
    L.nat_weights_noalloc.argtypes = [vp, vp, ctypes.c_int, vp, vp, vp, ctypes.c_int64, vp,
                                      ctypes.POINTER(ctypes.c_int64)]
    L.nat_spmm_csr.restype = ctypes.c_int
    L.nat_spmm_csr.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp,
                               ctypes.c_int, vp, ctypes.c_int, vp, vp, ctypes.c_int, vp]
    L.nat_weights_row_order.restype = ctypes.c_int
    L.nat_weights_row_order.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
    L.nat_csr_sort_rows.restype = ctypes.c_int
    L.nat_csr_sort_rows.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_int, vp]
    L.nat_topology_save.restype = ctypes.c_int
    L.nat_topology_save.argtypes = [vp, vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), vp]
    L.nat_grad_queries.restype = ctypes.c_int
    L.nat_grad_queries.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp, vp, vp, ctypes.c_int, vp,
                                   vp]
    L.nat_spmm_t_csr.restype = ctypes.c_int
    L.nat_spmm_t_csr.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp,
                                 vp, vp]
    L.nat_weights_spmm.restype = ctypes.c_int
    L.nat_weights_spmm.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, vp, vp]
    L.nat_topology_target.restype = ctypes.c_int
    L.nat_topology_target.argtypes = [vp, vp, ctypes.c_int64, ctypes.c_int]
    L.nat_topology_row_order.restype = ctypes.c_int
    L.nat_topology_row_order.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
    L.nat_batch_build.restype = ctypes.c_int
    L.nat_batch_build.argtypes = [ctypes.POINTER(dp), ip, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.POINTER(vp)]
    L.nat_batch_offsets.restype = ctypes.c_int
    L.nat_batch_offsets.argtypes = [vp, ctypes.POINTER(ctypes.c_int64),
                                    ctypes.POINTER(ctypes.c_int64)]
    L.nat_batch_weights_compute.restype = ctypes.c_int
    L.nat_batch_weights_compute.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_int64), vp,
                                            ctypes.POINTER(ctypes.c_int64)]
    L.nat_predicate_audit.restype = ctypes.c_int
    L.nat_predicate_audit.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_int64), vp]
    L.nat_graph_create.restype = ctypes.c_int
    L.nat_graph_create.argtypes = [vp, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64), vp,
                                   ctypes.POINTER(vp)]
    L.nat_graph_launch.restype = ctypes.c_int
    L.nat_graph_launch.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_int64)]
    L.nat_graph_copy_out.restype = ctypes.c_int
    L.nat_graph_copy_out.argtypes = [vp, vp, vp, vp, vp]
    L.nat_graph_free.restype = None
    L.nat_graph_free.argtypes = [vp]
    L.nat_insphere_classify.restype = ctypes.c_int
    L.nat_insphere_classify.argtypes = [dp, ctypes.c_int, dp, ip]
    L.nat_mesh_set_precision.restype = ctypes.c_int
    L.nat_mesh_set_precision.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.nat_mesh_set_predicates.restype = ctypes.c_int
    L.nat_mesh_set_predicates.argtypes = [vp, ctypes.c_int]
    L.nat_mesh_set_mixed_check.restype = ctypes.c_int
    L.nat_mesh_set_mixed_check.argtypes = [vp, ctypes.c_int]
    L.nat_mesh_build_ex.restype = ctypes.c_int
    L.nat_mesh_build_ex.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(vp)]
    L.nat_mesh_delaunay_audit.restype = ctypes.c_int
    L.nat_mesh_delaunay_audit.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.nat_predicate_exact.restype = ctypes.c_int
    L.nat_predicate_exact.argtypes = [dp, ctypes.c_int, ctypes.c_int, ip, ip]
    L.nat_mesh_from_packed.restype = ctypes.c_int
    L.nat_mesh_from_packed.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, vp, vp, vp, vp,
                                       vp, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    L.nat_mesh_set_policy.restype = ctypes.c_int
    L.nat_mesh_set_policy.argtypes = [vp, ctypes.POINTER(Policy), vp]
    L.nat_mesh_get_policy.restype = ctypes.c_int
    L.nat_mesh_get_policy.argtypes = [vp, ctypes.POINTER(Policy), vp]
    L.nat_mesh_autotune.restype = ctypes.c_int
    L.nat_mesh_autotune.argtypes = [vp, vp, ctypes.c_int64, vp, ctypes.c_int, vp,
                                    ctypes.POINTER(TuneReport)]
    L.nat_mesh_set_timing.restype = ctypes.c_int
    L.nat_mesh_set_timing.argtypes = [vp, ctypes.c_int]
    L.nat_mesh_get_timings.restype = ctypes.c_int
    L.nat_mesh_get_timings.argtypes = [vp, ctypes.POINTER(Timings)]
    L.nat_free_host.restype = None
    L.nat_free_host.argtypes = [vp]
    L.nat_last_error.restype = ctypes.c_char_p
    L.nat_last_error.argtypes = []
    _lib = L
    return L


def last_error():
    return lib().nat_last_error().decode(errors="replace")


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")
