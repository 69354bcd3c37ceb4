"""Seeded synthetic graphs from the library's generators (gen.hip).

Each returns a ``HostCSR`` of A_T (row v = in-neighbours of v, 1-based Int32
arrays, sorted, deduplicated) that converts to the storage types.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from ._lib import GGCsrHost, check, lib
from .storage import _CODE2NP, dtype_code

UNIT, UNIFORM01, INTEGER = 0, 1, 2


@dataclass
class HostCSR:
    m: int
    n: int
    rowptr: np.ndarray  # int32 [m+1], 1-based
    colval: np.ndarray  # int32 [nnz], 1-based
    nzval: np.ndarray
    symmetric: bool = False  # the pattern is symmetric (A_T rows are also out-neighbour lists)

    @property
    def nnz(self):
        return len(self.colval)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.nzval, self.colval - 1, self.rowptr.astype(np.int64) - 1),
                             shape=(self.m, self.n))

    def to_csr(self, backend):
        from .storage import SparseGPUMatrixCSR
        M = SparseGPUMatrixCSR(self.m, self.n, self.rowptr, self.colval, self.nzval, backend)
        M.symmetric = self.symmetric
        return M

    def to_sell(self, backend, slice_size=64, sigma=4096, heavy_threshold=0, heavy_seg_len=4096,
                relabel=False, sort_entries=False, rcm=False):
        from .storage import SparseGPUMatrixSELL
        M = SparseGPUMatrixSELL.__new__(SparseGPUMatrixSELL)
        M._build(self.m, self.n, self.rowptr, self.colval, np.ascontiguousarray(self.nzval),
                 backend, slice_size, sigma, heavy_threshold, heavy_seg_len, relabel, sort_entries,
                 rcm)
        M.symmetric = self.symmetric
        return M

    def degrees(self):
        return np.diff(self.rowptr.astype(np.int64))


def _take(out) -> HostCSR:
    try:
        h = out.contents
        dt = _CODE2NP[h.val_dtype]

        def arr(p, count, t):
            if count == 0:
                return np.zeros(0, t)
            nb = count * np.dtype(t).itemsize
            return np.frombuffer((C.c_char * nb).from_address(p), dtype=t).copy()

        return HostCSR(int(h.m), int(h.n), arr(h.rowptr, h.m + 1, np.int32),
                       arr(h.colval, h.nnz, np.int32),
                       arr(h.nzval, h.nnz, np.uint8 if dt == np.bool_ else dt).astype(dt))
    finally:
        lib().gg_csr_host_free(out)


def rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, permute=True, dtype=np.bool_,
         values=UNIT) -> HostCSR:
    """R-MAT graph (symmetrised, no self-loops or duplicates, ids permuted)."""
    out = C.POINTER(GGCsrHost)()
    check(lib().gg_gen_rmat(scale, edge_factor, a, b, c, seed, int(permute), dtype_code(dtype),
                            values, C.byref(out)), "rmat")
    H = _take(out)
    H.symmetric = True   # both directions of every edge
    return H


def grid(rows, cols, long_range_frac=0.01, wmin=1, wmax=1000, seed=1, dtype=np.float32) -> HostCSR:
    """4-neighbour grid plus random long-range edges; independent integer
    weights per direction in [wmin, wmax]."""
    out = C.POINTER(GGCsrHost)()
    check(lib().gg_gen_grid(rows, cols, long_range_frac, wmin, wmax, seed, dtype_code(dtype),
                            C.byref(out)), "grid")
    H = _take(out)
    H.symmetric = True   # both directions of every edge
    return H


def uniform_rows(n, per_row=16, seed=1, dtype=np.float32) -> HostCSR:
    """n x n, exactly per_row distinct uniform columns per row, values U[0,1)."""
    out = C.POINTER(GGCsrHost)()
    check(lib().gg_gen_uniform_rows(n, per_row, seed, dtype_code(dtype), C.byref(out)),
          "uniform_rows")
    return _take(out)


def random_graph(n, degree=16, seed=1, dtype=np.bool_, values=UNIT, wmin=1, wmax=1000) -> HostCSR:
    """n*degree/2 uniform vertex pairs, symmetrised."""
    out = C.POINTER(GGCsrHost)()
    check(lib().gg_gen_random_graph(n, degree, seed, dtype_code(dtype), values, wmin, wmax,
                                    C.byref(out)), "random_graph")
    H = _take(out)
    H.symmetric = True   # both directions of every edge
    return H


def read_matrix_market(path, dtype=None, transpose=False) -> HostCSR:
    """MatrixMarket coordinate file (real / integer / pattern; general,
    symmetric or skew-symmetric) as a HostCSR: the stored matrix, or its
    transpose; duplicates add up; symmetric files come back with both
    triangles and `symmetric` set.  dtype None keeps the file's type (pattern
    -> bool, integer -> int64, real -> float64) (gg_read_matrix_market)."""
    out = C.POINTER(GGCsrHost)()
    sym = C.c_int(0)
    code = 0 if dtype is None else dtype_code(dtype)
    check(lib().gg_read_matrix_market(os.fsencode(path), code, int(bool(transpose)), C.byref(out),
                                      C.byref(sym)), "read_matrix_market")
    H = _take(out)
    H.symmetric = bool(sym.value)
    return H


def read_edge_list(path, base=1, directed=False, dtype=np.bool_) -> HostCSR:
    """Edge list ("u v" or "u v w" per line, '#' / '%' comments, vertices
    counted from `base`) as the HostCSR of A_T (row v lists the tails of the
    edges into v); undirected lists get both directions and `symmetric`; a
    repeated edge keeps its smallest weight (gg_read_edge_list)."""
    out = C.POINTER(GGCsrHost)()
    check(lib().gg_read_edge_list(os.fsencode(path), int(base), int(bool(directed)),
                                  dtype_code(dtype), C.byref(out)), "read_edge_list")
    H = _take(out)
    H.symmetric = not directed
    return H


def pack(graphs):
    """Block-diagonal packing of square HostCSRs (A_T each, one value type):
    (packed HostCSR, offsets) where graph g owns rows / columns
    offsets[g] .. offsets[g+1]-1 (0-based), each block padded with isolated
    vertices to a multiple of 256 rows (gg_pack_graphs)."""
    hs = []
    for g in graphs:
        nz = np.ascontiguousarray(g.nzval)
        if nz.dtype == np.bool_:
            nz = nz.view(np.uint8)
        h = GGCsrHost()
        h.m, h.n, h.nnz = g.m, g.n, g.nnz
        h.val_dtype = dtype_code(g.nzval.dtype)
        rp = np.ascontiguousarray(g.rowptr, np.int32)
        cv = np.ascontiguousarray(g.colval, np.int32)
        h.rowptr, h.colval, h.nzval = rp.ctypes.data, cv.ctypes.data, nz.ctypes.data
        hs.append((h, rp, cv, nz))   # the arrays stay alive through the call
    arr = (C.POINTER(GGCsrHost) * len(hs))(*[C.pointer(h[0]) for h in hs])
    offsets = np.zeros(len(hs) + 1, np.int64)
    out = C.POINTER(GGCsrHost)()
    check(lib().gg_pack_graphs(len(hs), arr, C.byref(out), offsets.ctypes.data), "pack")
    P = _take(out)
    # a block-diagonal pack of symmetric graphs is symmetric (top-down steps)
    P.symmetric = all(getattr(g, "symmetric", False) for g in graphs)
    return P, offsets


def rcm_order(H):
    """Reverse Cuthill-McKee order of a square HostCSR pattern (gg_rcm_order):
    1-based vertices, position by position."""
    if H.m != H.n:
        raise ValueError("rcm_order needs a square matrix")
    out = np.zeros(H.m, np.int32)
    rp = np.ascontiguousarray(H.rowptr, np.int32)
    cv = np.ascontiguousarray(H.colval, np.int32)
    check(lib().gg_rcm_order(H.m, rp.ctypes.data, cv.ctypes.data, out.ctypes.data), "rcm_order")
    return out
