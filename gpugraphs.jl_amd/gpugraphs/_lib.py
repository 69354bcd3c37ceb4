"""ctypes binding of libgpugraphs_hip.so (the C ABI of include/gpugraphs.h).

The library is loaded from ``gpugraphs.jl_amd/lib``; a missing library is an
error, never a silent fallback.  Status codes become the exception types the
reference throws for the same conditions (see ``errors``).
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import sys
import threading
import weakref

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libgpugraphs_hip.so")

# gg_dtype
BOOL, INT32, INT64, FLOAT32, FLOAT64, UINT64 = 1, 2, 3, 4, 5, 6
# gg_format
CSR, SELL, CSC = 1, 2, 3
# gg_layout
COL_MAJOR, ROW_MAJOR = 0, 1
FLAG_UNIT_PATTERN = 1
FLAG_INTERNAL_IDS = 2
FLAG_SYMMETRIC = 4
SELL_RELABEL = 1
SELL_SORT_ENTRIES = 2
SELL_RCM = 4


class DimensionMismatch(ValueError):
    """Julia DimensionMismatch."""


class ArgumentError(ValueError):
    """Julia ArgumentError."""


class BoundsError(IndexError):
    """Julia BoundsError."""


class MethodError(TypeError):
    """No kernel for this operator / element-type combination."""


class HIPError(RuntimeError):
    """A HIP runtime call failed."""


class NotConvergedError(RuntimeError):
    """Bellman-Ford reached its iteration cap (negative cycle)."""


class BudgetExceeded(RuntimeError):
    """A request does not fit the serving memory budget."""


_STATUS_EXC = {1: DimensionMismatch, 2: ArgumentError, 3: AssertionError, 4: BoundsError,
               5: MethodError, 6: HIPError, 7: MemoryError, 8: NotConvergedError,
               9: BudgetExceeded}


class GGMatrix(C.Structure):
    _fields_ = [
        ("format", C.c_int32), ("val_dtype", C.c_int32),
        ("m", C.c_int64), ("n", C.c_int64), ("nnz", C.c_int64),
        ("ptr", C.c_void_p), ("idx", C.c_void_p), ("val", C.c_void_p),
        ("slice_size", C.c_int64), ("nslices", C.c_int64), ("n_stored", C.c_int64),
        ("perm", C.c_void_p), ("slot_len", C.c_void_p),
        ("n_heavy", C.c_int64), ("heavy_row", C.c_void_p), ("heavy_ptr", C.c_void_p),
        ("heavy_seg", C.c_void_p), ("heavy_seg_len", C.c_int64), ("heavy_nseg", C.c_int64),
        ("heavy_col", C.c_void_p), ("heavy_val", C.c_void_p), ("flags", C.c_int64),
        ("vinv", C.c_void_p),
    ]


class GGSellHost(C.Structure):
    _fields_ = [
        ("m", C.c_int64), ("n", C.c_int64), ("nnz", C.c_int64), ("slice_size", C.c_int64),
        ("sigma", C.c_int64), ("nslices", C.c_int64), ("n_stored", C.c_int64),
        ("n_slots", C.c_int64), ("val_dtype", C.c_int32),
        ("slice_ptr", C.c_void_p), ("colval", C.c_void_p), ("nzval", C.c_void_p),
        ("perm", C.c_void_p), ("slot_len", C.c_void_p),
        ("n_heavy", C.c_int64), ("heavy_nnz", C.c_int64), ("heavy_seg_len", C.c_int64),
        ("n_seg", C.c_int64),
        ("heavy_row", C.c_void_p), ("heavy_ptr", C.c_void_p), ("heavy_seg", C.c_void_p),
        ("heavy_col", C.c_void_p), ("heavy_val", C.c_void_p),
        ("vinv", C.c_void_p), ("n_internal", C.c_int64),
    ]


class GGSellSizes(C.Structure):
    _fields_ = [("slice_size", C.c_int64), ("nslices", C.c_int64), ("n_stored", C.c_int64),
                ("n_slots", C.c_int64), ("n_heavy", C.c_int64), ("heavy_nnz", C.c_int64),
                ("n_seg", C.c_int64), ("n_internal", C.c_int64), ("has_perm", C.c_int32),
                ("has_vinv", C.c_int32)]


class GGCsrHost(C.Structure):
    _fields_ = [("m", C.c_int64), ("n", C.c_int64), ("nnz", C.c_int64),
                ("val_dtype", C.c_int32), ("rowptr", C.c_void_p), ("colval", C.c_void_p),
                ("nzval", C.c_void_p)]


class QueryInfo(C.Structure):
    _fields_ = [("request_id", C.c_int64), ("graph", C.c_int32), ("kind", C.c_int32),
                ("source", C.c_int64), ("status", C.c_int32), ("arrival_s", C.c_double),
                ("admit_s", C.c_double), ("complete_s", C.c_double)]


class ServerCounters(C.Structure):
    _fields_ = [("completed", C.c_int64), ("failed", C.c_int64), ("waiting", C.c_int64),
                ("running", C.c_int64), ("steps", C.c_int64), ("arena_bytes", C.c_size_t),
                ("arena_used", C.c_size_t)]


class EngineCounters(C.Structure):
    _fields_ = [("steps", C.c_int64), ("busy_slot_steps", C.c_int64), ("finished", C.c_int64),
                ("slots", C.c_int32), ("reserved", C.c_int32), ("device_s", C.c_double)]


class TuneCandidate(C.Structure):
    _fields_ = [("format", C.c_int32), ("csr_lanes", C.c_int32), ("slice_size", C.c_int64),
                ("sigma", C.c_int64), ("heavy_threshold", C.c_int64), ("heavy_seg_len", C.c_int64),
                ("sell_flags", C.c_int32), ("reserved", C.c_int32), ("ms", C.c_double)]


# every symbol of include/gpugraphs.h with its ctypes signature
_vp, _i64, _i32, _u64, _d = C.c_void_p, C.c_int64, C.c_int, C.c_uint64, C.c_double
_PM = C.POINTER(GGMatrix)
SIGNATURES = {
    "gg_abi_version": ([], _i32),
    "gg_status_string": ([_i32], C.c_char_p),
    "gg_last_error": ([], C.c_char_p),
    "gg_ctx_create": ([_i32, _vp, C.POINTER(_vp)], _i32),
    "gg_ctx_destroy": ([_vp], _i32),
    "gg_ctx_synchronize": ([_vp], _i32),
    "gg_ctx_set_stream": ([_vp, _vp], _i32),
    "gg_malloc": ([_vp, C.c_size_t, C.POINTER(_vp)], _i32),
    "gg_free": ([_vp, _vp], _i32),
    "gg_memcpy_htod": ([_vp, _vp, _vp, C.c_size_t], _i32),
    "gg_memcpy_dtoh": ([_vp, _vp, _vp, C.c_size_t], _i32),
    "gg_ctx_enable_timing": ([_vp, _i32], _i32),
    "gg_ctx_set_graph_steps": ([_vp, _i32], _i32),
    "gg_ctx_set_csr_lanes": ([_vp, _i32], _i32),
    "gg_ctx_set_push_threshold": ([_vp, _d], _i32),
    "gg_ctx_set_sssp_threshold": ([_vp, _d], _i32),
    "gg_ctx_last_kernel_ms": ([_vp, C.POINTER(_d)], _i32),
    "gg_ctx_hot_kernel_stats": ([_vp, C.POINTER(_d), C.POINTER(_i64), _i32], _i32),
    "gg_matrix_check": ([_PM], _i32),
    "gg_matrix_slots": ([_PM, C.POINTER(_i64)], _i32),
    "gg_csr_check": ([_i64, _i64, _vp, _i64, _vp, _i64, _i64], _i32),
    "gg_sell_check": ([_i64, _i64, _i64, _i64, _vp, _i64, _i64], _i32),
    "gg_rcm_order": ([_i64, _vp, _vp, _vp], _i32),
    "gg_sell_build": ([_i64, _i64, _vp, _vp, _vp, _i32, _i64, _i64, _i64, _i64, _i32,
                       C.POINTER(C.POINTER(GGSellHost))], _i32),
    "gg_sell_host_free": ([C.POINTER(GGSellHost)], None),
    "gg_sell_plan_device": ([_vp, _i64, _i64, _vp, _vp, _vp, _i32, _i64, _i64, _i64, _i64, _i32,
                             C.POINTER(_vp), C.POINTER(GGSellSizes)], _i32),
    "gg_sell_fill_device": ([_vp, _PM], _i32),
    "gg_sell_plan_free": ([_vp], None),
    "gg_csr_host_free": ([C.POINTER(GGCsrHost)], None),
    "gg_gen_rmat": ([_i32, _i32, _d, _d, _d, _u64, _i32, _i32, _i32,
                     C.POINTER(C.POINTER(GGCsrHost))], _i32),
    "gg_gen_grid": ([_i64, _i64, _d, C.c_int32, C.c_int32, _u64, _i32,
                     C.POINTER(C.POINTER(GGCsrHost))], _i32),
    "gg_gen_uniform_rows": ([_i64, _i64, _u64, _i32, C.POINTER(C.POINTER(GGCsrHost))], _i32),
    "gg_pack_graphs": ([_i32, _vp, C.POINTER(C.POINTER(GGCsrHost)), _vp], _i32),
    "gg_read_matrix_market": ([C.c_char_p, _i32, _i32, C.POINTER(C.POINTER(GGCsrHost)), _vp],
                              _i32),
    "gg_read_edge_list": ([C.c_char_p, _i32, _i32, _i32, C.POINTER(C.POINTER(GGCsrHost))], _i32),
    "gg_bfs_packed": ([_vp, _PM, _vp, _i32, _vp, _vp, _vp], _i32),
    "gg_bfs_packed_do": ([_vp, _PM, _PM, _vp, _i32, _vp, _vp, _vp, _d, _d], _i32),
    "gg_gen_random_graph": ([_i64, _i64, _u64, _i32, _i32, C.c_int32, C.c_int32,
                             C.POINTER(C.POINTER(GGCsrHost))], _i32),
    "gg_spmv": ([_vp, _vp, _i32, _i64, _PM, _vp, _i32, _i64, _i32, _i32, _i32, _vp, _i32,
                 _i64, _i64, _i64], _i32),
    "gg_spmv_rows": ([_vp, _vp, _i32, _i64, _PM, _vp, _i32, _i64, _i32, _i32, _i32, _vp, _i64],
                     _i32),
    "gg_spmm": ([_vp, _vp, _i32, _i64, _i64, _i32, _PM, _vp, _i32, _i64, _i64, _i32, _i32,
                 _i32, _i32, _vp, _i32, _i64, _i64, _i64], _i32),
    "gg_spmspv": ([_vp, _vp, _vp, C.POINTER(_i64), _i32, _i64, _PM, _vp, _vp, _i64, _i32, _i64,
                   _i32, _i32, _vp, _i32, _i64, _i32], _i32),
    "gg_bfs_distances": ([_vp, _PM, _i64, _vp, _i32], _i32),
    "gg_bfs_distances_noalloc": ([_vp, _PM, _i64, _vp, _i32], _i32),
    "gg_bfs_parents": ([_vp, _PM, _i64, _vp, _i64], _i32),
    "gg_bfs_distances_do": ([_vp, _PM, _PM, _i64, _vp, _d, _d], _i32),
    "gg_shortest_path_delta": ([_vp, _PM, _i64, _vp, _i32, _d, _vp], _i32),
    "gg_shortest_path_delta_multi": ([_vp, _PM, _vp, _i64, _vp, _i32, _d, _vp], _i32),
    "gg_shortest_path": ([_vp, _PM, _i64, _vp, _i32, C.POINTER(_i64)], _i32),
    "gg_shortest_path_multi": ([_vp, _PM, _vp, _i64, _vp, _i32, C.POINTER(_i64)], _i32),
    "gg_msbfs_distances": ([_vp, _PM, _vp, _i64, _vp, C.POINTER(_i64)], _i32),
    "gg_mssp_distances": ([_vp, _PM, _vp, _i64, _vp, _i32, C.POINTER(_i64)], _i32),
    "gg_tune_default_candidates": ([_PM, _i64, C.POINTER(TuneCandidate), _i64, C.POINTER(_i64)],
                                   _i32),
    "gg_tune_layout": ([_vp, _PM, _i32, _vp, _i64, C.POINTER(TuneCandidate), _i64, _i32,
                        C.POINTER(TuneCandidate)], _i32),
    "gg_tune_push_threshold": ([_vp, _PM, _vp, _i64, _vp, _vp, _i64, _i32, C.POINTER(_d)], _i32),
    "gg_tune_sssp_threshold": ([_vp, _PM, _vp, _i64, _vp, _vp, _i64, _i32, C.POINTER(_d)], _i32),
    "gg_server_create": ([_i32, C.c_size_t, C.POINTER(_vp)], _i32),
    "gg_server_add_graph": ([_vp, _PM, C.POINTER(_i32)], _i32),
    "gg_server_start": ([_vp], _i32),
    "gg_server_set_pacing": ([_vp, _d], _i32),
    "gg_server_submit": ([_vp, _i32, _i32, _i64, C.POINTER(_i64)], _i32),
    "gg_server_poll": ([_vp, C.POINTER(QueryInfo), _i64, C.POINTER(_i64), _d], _i32),
    "gg_server_fetch": ([_vp, _i64, _vp, C.c_size_t, _d], _i32),
    "gg_server_fetch_device": ([_vp, _i64, _vp, C.c_size_t, _d], _i32),
    "gg_server_release": ([_vp, _i64, _d], _i32),
    "gg_server_stats": ([_vp, C.POINTER(ServerCounters)], _i32),
    "gg_server_engine_stats": ([_vp, _i32, _i32, C.POINTER(EngineCounters)], _i32),
    "gg_server_stop": ([_vp], _i32),
    "gg_server_destroy": ([_vp], _i32),
}

_lib = None
_lock = threading.Lock()


def lib():
    """The loaded library (raises when it has not been built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(
                        f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                        "(make -C gpugraphs.jl_amd/csrc)")
                L = C.CDLL(LIB_PATH)
                for name, (args, res) in SIGNATURES.items():
                    f = getattr(L, name)
                    f.argtypes = args
                    f.restype = res
                _lib = L
    return _lib


def check(status: int, what: str = ""):
    """Raise the exception that matches a gg_status."""
    if status == 0:
        return
    msg = lib().gg_last_error().decode(errors="replace")
    exc = _STATUS_EXC.get(status, RuntimeError)
    raise exc(f"{what}: {msg}" if what else msg)


# ------------------------------------------------------------------ contexts
_ctxs: dict = {}
_servers = weakref.WeakSet()   # live GraphServers, closed before the contexts at exit
_teardown_registered = False


def shutdown():
    """Explicit teardown, run from an atexit hook registered after torch
    initialised the GPU (so it runs before torch's own hooks and before the
    HIP runtime unloads): stop and destroy every live server first (their
    worker threads step engines on the device), then every cached context
    with its captured graphs, events, pinned buffer and workspace.  No HIP
    object of this library outlives the runtime."""
    for srv in list(_servers):
        try:
            srv.close()
        except Exception:
            pass
    while _ctxs:
        _, ctx = _ctxs.popitem()
        lib().gg_ctx_destroy(ctx)


def _register_teardown():
    global _teardown_registered
    if not _teardown_registered:
        atexit.register(shutdown)
        _teardown_registered = True


def register_server(srv):
    """Track a server so that shutdown() closes it before the contexts go."""
    _register_teardown()
    _servers.add(srv)


def finalizing(_is_finalizing=sys.is_finalizing) -> bool:
    """True once the interpreter is shutting down (no native calls then).
    sys.is_finalizing is bound at definition time, so this still answers
    when module globals have been cleared."""
    return _is_finalizing()


def context(device: int = 0, stream=None):
    """gg_ctx bound to (device, stream); the torch current stream by default."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("GPUGraphs kernels need a ROCm GPU; none is visible")
    if stream is None:
        stream = torch.cuda.current_stream(device)
    key = (device, int(stream.cuda_stream))
    ctx = _ctxs.get(key)
    if ctx is None:
        _register_teardown()
        out = C.c_void_p()
        check(lib().gg_ctx_create(device, C.c_void_p(key[1]), C.byref(out)), "gg_ctx_create")
        ctx = out
        _ctxs[key] = ctx
    return ctx


def last_kernel_ms(ctx) -> float:
    ms = C.c_double()
    check(lib().gg_ctx_last_kernel_ms(ctx, C.byref(ms)))
    return ms.value


def enable_timing(ctx, on: bool = True):
    check(lib().gg_ctx_enable_timing(ctx, int(on)))


def set_graph_steps(ctx, steps: int):
    """Run level / iteration loops as HIP graphs of `steps` captured steps
    (0: one host-launched step at a time)."""
    check(lib().gg_ctx_set_graph_steps(ctx, int(steps)))


def set_csr_lanes(ctx, lanes: int):
    """Lanes per CSR row (power of two in 1:64; 0 = by average row length)."""
    check(lib().gg_ctx_set_csr_lanes(ctx, int(lanes)))


def set_push_threshold(ctx, alpha: float):
    """Multi-source BFS levels run top-down when frontier out-edges * alpha <
    entries (0: never)."""
    check(lib().gg_ctx_set_push_threshold(ctx, float(alpha)))


def set_sssp_threshold(ctx, alpha: float):
    """k-source Bellman-Ford iterations on a symmetric relabelled SELL relax
    only the rows next to the fallen rows when those rows' out-edges * alpha <
    entries (0: never)."""
    check(lib().gg_ctx_set_sssp_threshold(ctx, float(alpha)))


def hot_kernel_stats(ctx, reset: bool = True):
    """(summed ms, launches) of hot-kernel launches since the last reset."""
    ms, n = C.c_double(), C.c_int64()
    check(lib().gg_ctx_hot_kernel_stats(ctx, C.byref(ms), C.byref(n), int(reset)))
    return ms.value, n.value
