"""BFS and Bellman-Ford (src/algorithms/bfs.jl, src/algorithms/shortest_path.jl).

``A_T`` is the transposed adjacency (row v lists the in-neighbours of v), as
in the reference (test/bfs.jl:24).  Sources are 1-based.  Results are device
int32 / value-type tensors with the reference's sentinels: unreached =
typemax, the source's parent = 0.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import ArgumentError, check, context, lib
from .storage import (ROCBackend, SparseGPUMatrixCSR, _torch_dtype, dtype_code, get_backend,
                      np_dtype_of, ptr)


def _dev(A):
    b = get_backend(A)
    if not isinstance(b, ROCBackend):
        raise RuntimeError("GPUGraphs kernels need a ROCBackend matrix; got " + repr(b))
    return b.device


def _empty(n, dtype, device):
    import torch
    return torch.empty(max(n, 0), dtype=_torch_dtype(dtype), device=f"cuda:{device}")


def bfs_distances_(A_T, source, dist, curr=None, next=None, to_explore=None, use_mask=True):
    """bfs_distances! (src/algorithms/bfs.jl:31-73): levels into ``dist``.

    ``curr`` / ``next`` / ``to_explore`` are accepted for signature parity; the
    level kernel keeps its frontier as bitmaps in the context workspace."""
    dev = _dev(A_T)
    d = A_T.descriptor()
    check(lib().gg_bfs_distances(context(dev), C.byref(d), int(source), ptr(dist), int(use_mask)),
          "bfs_distances")
    return None


def bfs_distances(A_T, source, use_mask=True):
    """bfs_distances (src/algorithms/bfs.jl:1-29)."""
    dist = _empty(A_T.m, np.int32, _dev(A_T))
    bfs_distances_(A_T, source, dist, use_mask=use_mask)
    return dist


def bfs_packed(A_T, offsets, sources):
    """One BFS per graph of a block-diagonal pack (generators.pack) in one
    persistent launch: (dist over all packed rows, levels per graph);
    dist[offsets[g]:offsets[g+1]][:n_g] equals bfs_distances(graph g,
    sources[g])."""
    if not isinstance(A_T, SparseGPUMatrixCSR):
        raise ArgumentError("a packed graph is a SparseGPUMatrixCSR")
    dev = _dev(A_T)
    off = np.ascontiguousarray(offsets, np.int64)
    src = np.ascontiguousarray(sources, np.int32)
    if len(off) != len(src) + 1:
        raise ArgumentError("offsets needs one entry more than sources")
    dist = _empty(A_T.m, np.int32, dev)
    levels = np.zeros(len(src), np.int32)
    d = A_T.descriptor()
    check(lib().gg_bfs_packed(context(dev), C.byref(d), off.ctypes.data, len(src),
                              src.ctypes.data, ptr(dist), levels.ctypes.data), "bfs_packed")
    return dist, levels


def bfs_packed_do(A_T, offsets, sources, A_out=None, alpha=-1.0, beta=-1.0):
    """The results of bfs_packed from one direction-optimising traversal of
    the pack seeded with every graph's source: (dist over all packed rows,
    levels per graph).  A_T may be any CSR or SELL layout of the pack; A_out
    is the pack's CSR of A (for a symmetric pack, the packed CSR; the default
    when A_T is that CSR)."""
    dev = _dev(A_T)
    if A_out is None:
        if not isinstance(A_T, SparseGPUMatrixCSR):
            raise ArgumentError("A_out (the packed CSR of A) is needed with a SELL A_T")
        A_out = A_T
    if not isinstance(A_out, SparseGPUMatrixCSR):
        raise ArgumentError("A_out must be a SparseGPUMatrixCSR")
    off = np.ascontiguousarray(offsets, np.int64)
    src = np.ascontiguousarray(sources, np.int32)
    if len(off) != len(src) + 1:
        raise ArgumentError("offsets needs one entry more than sources")
    dist = _empty(A_T.m, np.int32, dev)
    levels = np.zeros(len(src), np.int32)
    d, o = A_T.descriptor(), A_out.descriptor()
    check(lib().gg_bfs_packed_do(context(dev), C.byref(d), C.byref(o), off.ctypes.data, len(src),
                                 src.ctypes.data, ptr(dist), levels.ctypes.data, float(alpha),
                                 float(beta)), "bfs_packed_do")
    return dist, levels


def bfs_distances_do(A_T, source, A_out=None, alpha=-1.0, beta=-1.0):
    """Direction-optimising BFS levels (push over A_out's rows, pull over A_T).
    A_out = CSR of A; for a symmetric graph it is A_T itself (the default)."""
    dev = _dev(A_T)
    if A_out is None:
        A_out = A_T
    if not isinstance(A_out, SparseGPUMatrixCSR):
        raise ArgumentError("A_out must be a SparseGPUMatrixCSR")
    dist = _empty(A_T.m, np.int32, dev)
    d, o = A_T.descriptor(), A_out.descriptor()
    check(lib().gg_bfs_distances_do(context(dev), C.byref(d), C.byref(o), int(source), ptr(dist),
                                    float(alpha), float(beta)), "bfs_distances_do")
    return dist


def bfs_parents_(A_T, source, parents, curr=None, next=None, to_explore=None, batch_size=8):
    """bfs_parents! (src/algorithms/bfs.jl:135-183)."""
    dev = _dev(A_T)
    d = A_T.descriptor()
    check(lib().gg_bfs_parents(context(dev), C.byref(d), int(source), ptr(parents),
                               int(batch_size)), "bfs_parents")
    return None


def bfs_parents(A_T, source, batch_size=8):
    """bfs_parents (src/algorithms/bfs.jl:113-133)."""
    par = _empty(A_T.m, np.int32, _dev(A_T))
    bfs_parents_(A_T, source, par, batch_size=batch_size)
    return par


def dist_dtype(A_T):
    """T_dist = ifelse(Tv == Bool, Int32, Tv), src/algorithms/shortest_path.jl:18."""
    vt = np_dtype_of(A_T.nzval)
    return np.dtype(np.int32) if vt == np.bool_ else vt


def shortest_path_(A_T, source, dist, next=None, updated=None, diff=None):
    """shortest_path! for one source (Int) or many (a sequence of Ints)."""
    dev = _dev(A_T)
    d = A_T.descriptor()
    it = C.c_int64()
    dt = dtype_code(np_dtype_of(dist))
    if np.ndim(source) == 0:
        check(lib().gg_shortest_path(context(dev), C.byref(d), int(source), ptr(dist), dt,
                                     C.byref(it)), "shortest_path")
    else:
        src = np.ascontiguousarray(source, np.int32)
        check(lib().gg_shortest_path_multi(context(dev), C.byref(d), src.ctypes.data, len(src),
                                           ptr(dist), dt, C.byref(it)), "shortest_path")
    return it.value


def shortest_path(A_T, source):
    """shortest_path (src/algorithms/shortest_path.jl:12-32 and :88-110).

    One source: a length-m vector.  A vector of k sources: an m x k
    column-major (Julia layout) matrix, column i from sources[i]."""
    dev = _dev(A_T)
    dt = dist_dtype(A_T)
    if np.ndim(source) == 0:
        dist = _empty(A_T.m, dt, dev)
    else:
        k = len(source)
        dist = _empty(A_T.m * k, dt, dev).view(k, A_T.m).T  # column-major m x k
    shortest_path_(A_T, source, dist)
    return dist


def shortest_path_delta(A_out, source, delta=0.0, return_phases=False):
    """The distances of shortest_path by delta-stepping: push over A_out, the
    CSR of A (row u = out-neighbours of u with the weights of u -> v; for a
    symmetric matrix, A_T itself).  delta = 0 takes 4 x the mean weight.
    Weights reached from the source must be non-negative.  One source: a
    length-m vector; a sequence of k sources: an m x k column-major matrix
    whose columns run together (shared phases)."""
    if not isinstance(A_out, SparseGPUMatrixCSR):
        raise ArgumentError("delta-stepping runs on a SparseGPUMatrixCSR of A")
    dev = _dev(A_out)
    dt = dist_dtype(A_out)
    d = A_out.descriptor()
    ph = C.c_int64()
    if np.ndim(source) == 0:
        dist = _empty(A_out.m, dt, dev)
        check(lib().gg_shortest_path_delta(context(dev), C.byref(d), int(source), ptr(dist),
                                           dtype_code(dt), float(delta), C.byref(ph)),
              "shortest_path_delta")
    else:
        src = np.ascontiguousarray(source, np.int32)
        k = len(src)
        dist = _empty(A_out.m * k, dt, dev).view(k, A_out.m).T  # column-major m x k
        check(lib().gg_shortest_path_delta_multi(context(dev), C.byref(d), src.ctypes.data, k,
                                                 ptr(dist), dtype_code(dt), float(delta),
                                                 C.byref(ph)), "shortest_path_delta")
    return (dist, ph.value) if return_phases else dist


def multi_source_bfs(A_T, sources):
    """Levels from k sources at once (or-and SpMM over 64-bit source masks);
    an m x k column-major int32 matrix whose column i equals
    bfs_distances(A_T, sources[i])."""
    dev = _dev(A_T)
    src = np.ascontiguousarray(sources, np.int32)
    k = len(src)
    dist = _empty(A_T.m * k, np.int32, dev)
    d = A_T.descriptor()
    lv = C.c_int64()
    check(lib().gg_msbfs_distances(context(dev), C.byref(d), src.ctypes.data, k, ptr(dist),
                                   C.byref(lv)), "multi_source_bfs")
    return dist.view(k, A_T.m).T
