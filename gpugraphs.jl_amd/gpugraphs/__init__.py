"""GPUGraphs on MI355X: semiring sparse kernels and graph loops in HIP.

Mirrors the export list of the reference module (src/GPUGraphs.jl:25-51):
storage types, gpu_spmv! / gpu_spmm! (as gpu_spmv_ / gpu_spmm_), sprand_gpu,
monoid_neutral / monoid_absorb, bfs_distances[!], bfs_parents[!],
shortest_path[!] and the GPUGraphs_* operators.  The kernels are in
libgpugraphs_hip.so (C ABI: include/gpugraphs.h).
"""
from ._lib import (ArgumentError, BoundsError, BudgetExceeded, DimensionMismatch, HIPError,
                   MethodError, NotConvergedError, lib)
from .algebra import (GPUGraphs_add, GPUGraphs_any, GPUGraphs_band, GPUGraphs_bor, GPUGraphs_first,
                      GPUGraphs_land, GPUGraphs_lor, GPUGraphs_max, GPUGraphs_min, GPUGraphs_mul,
                      GPUGraphs_pair, GPUGraphs_safe_add, GPUGraphs_second, GPUGraphs_secondi,
                      monoid_absorb, monoid_neutral)
from .algorithms import (bfs_distances, bfs_distances_, bfs_distances_do, bfs_packed, bfs_packed_do,
                         bfs_parents, bfs_parents_, dist_dtype, multi_source_bfs, shortest_path,
                         shortest_path_, shortest_path_delta)
from .ops import csc_view, gpu_spmm_, gpu_spmspv, gpu_spmspv_, gpu_spmv_
from .storage import (AbstractSparseGPUMatrix, CPUBackend, ROCBackend, SparseGPUMatrixCSC,
                      SparseGPUMatrixCSR, SparseGPUMatrixSELL, SparseGPUVector, get_backend,
                      sprand_gpu)

AbstractSparseGPUVector = SparseGPUVector

__all__ = [
    "AbstractSparseGPUMatrix", "SparseGPUMatrixCSR", "SparseGPUMatrixCSC", "SparseGPUMatrixSELL",
    "SparseGPUVector", "AbstractSparseGPUVector", "gpu_spmv_", "gpu_spmm_", "sprand_gpu",
    "monoid_neutral", "monoid_absorb", "bfs_distances", "bfs_distances_", "bfs_parents",
    "bfs_parents_", "shortest_path", "shortest_path_", "GPUGraphs_add", "GPUGraphs_safe_add",
    "GPUGraphs_mul", "GPUGraphs_second", "GPUGraphs_band", "GPUGraphs_bor", "GPUGraphs_any",
    "GPUGraphs_secondi", "GPUGraphs_max", "GPUGraphs_min",
    # beyond the reference's exports
    "GPUGraphs_first", "GPUGraphs_pair", "GPUGraphs_lor", "GPUGraphs_land", "gpu_spmspv",
    "gpu_spmspv_", "csc_view", "bfs_distances_do", "multi_source_bfs", "bfs_packed", "bfs_packed_do", "shortest_path_delta",
    "dist_dtype",
    "ROCBackend", "CPUBackend", "get_backend", "DimensionMismatch", "ArgumentError",
    "BoundsError", "MethodError", "HIPError", "NotConvergedError", "BudgetExceeded", "lib",
]
