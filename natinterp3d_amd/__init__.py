"""natinterp3d_amd -- natural neighbour (Sibson) interpolation in 3D with the
query path on AMD Instinct MI355X (gfx950).

Drop-in for natinterp3d's API (natinterp3d/__init__.py:7): ``Interpolator``,
``interpolate`` and ``get_weights`` return the same results (bit-identical
CSR matrices) with the weights computed by hand-written HIP kernels.
"""
from .batch import BatchInterpolator
from .natinterp3d import (GraphCapacityError, GraphPrecisionError, Interpolator,
                          MeshAndVerticesHIP, WeightsGraph, clear_mesh_cache, get_weights,
                          interpolate, mesh_cache_stats, spmm_device)

__all__ = ["BatchInterpolator", "GraphCapacityError", "GraphPrecisionError", "Interpolator",
           "MeshAndVerticesHIP",
           "WeightsGraph", "clear_mesh_cache", "get_weights", "interpolate", "mesh_cache_stats",
           "spmm_device"]
__version__ = "0.1.0"
