"""ctypes front of the CPU oracle (oracle/gg_oracle.cpp).

TEST INFRASTRUCTURE ONLY.  Tests, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this module as the checker; the
product package never does.

Arrays are numpy, 1-based Int32 indices exactly as the reference's structs
hold them (src/storage/CSR.jl:14-24, src/storage/SELL.jl:3-18).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

# codes shared with include/gpugraphs.h
BOOL, INT32, INT64, FLOAT32, FLOAT64 = 1, 2, 3, 4, 5
(OP_FIRST, OP_SECOND, OP_ANY, OP_PAIR, OP_ADD, OP_SAFE_ADD, OP_MUL, OP_MIN, OP_MAX,
 OP_LOR, OP_LAND, OP_BOR, OP_BAND, OP_SECONDI) = range(1, 15)
CSR, SELL, CSC = 1, 2, 3

STATUS = {0: "OK", 1: "DimensionMismatch", 2: "ArgumentError", 3: "AssertionError",
          4: "BoundsError", 5: "Unsupported", 8: "NotConverged"}

_NP2CODE = {np.dtype(np.bool_): BOOL, np.dtype(np.uint8): BOOL, np.dtype(np.int32): INT32,
            np.dtype(np.int64): INT64, np.dtype(np.float32): FLOAT32,
            np.dtype(np.float64): FLOAT64}
_CODE2NP = {BOOL: np.uint8, INT32: np.int32, INT64: np.int64, FLOAT32: np.float32,
            FLOAT64: np.float64}


class OracleError(RuntimeError):
    def __init__(self, status: int):
        super().__init__(STATUS.get(status, f"status {status}"))
        self.status = status


class _Mat(C.Structure):
    _fields_ = [("format", C.c_int32), ("val_dtype", C.c_int32), ("m", C.c_int64),
                ("n", C.c_int64), ("nnz", C.c_int64), ("ptr", C.c_void_p), ("idx", C.c_void_p),
                ("val", C.c_void_p), ("slice_size", C.c_int64), ("nslices", C.c_int64),
                ("n_stored", C.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle not built: run make -C {_HERE}")
        L = C.CDLL(_LIB_PATH)
        vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int
        L.orc_num_threads.restype = i32
        L.orc_csr_check.argtypes = [i64, i64, vp, i64, vp, i64, i64]
        L.orc_sell_size.argtypes = [i64, vp, i64, vp, vp, vp]
        L.orc_sell_fill.argtypes = [i64, vp, vp, vp, i32, i64, i64, vp, vp, vp]
        L.orc_spmv.argtypes = [vp, i32, i64, vp, vp, i32, i64, i32, i32, i32, vp, i32, i64, i64,
                               i64, i32]
        L.orc_spmm.argtypes = [vp, i32, i64, i64, vp, vp, i32, i64, i64, i32, i32, i32, vp, i32,
                               i64, i64, i64, i32]
        L.orc_bfs_distances.argtypes = [vp, i64, vp, i32, i32, vp]
        L.orc_bfs_parents.argtypes = [vp, i64, vp, i64, i32]
        L.orc_shortest_path.argtypes = [vp, i64, vp, i32, i32, i64, vp]
        L.orc_shortest_path_multi.argtypes = [vp, vp, i64, vp, i32, i32, i64, vp]
        L.orc_spmspv.argtypes = [vp, vp, vp, i32, i64, vp, vp, vp, i64, i32, i64, i32, i32, vp,
                                 i32, i64, i32]
        L.orc_bfs_queue_multi.argtypes = [vp, vp, i64, vp, i32]
        L.orc_sssp_certificate.argtypes = [vp, vp, i64, vp, i32, i32, vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def dcode(a: np.ndarray) -> int:
    return _NP2CODE[a.dtype]


def _check(s: int):
    if s != 0:
        raise OracleError(s)


@dataclass
class Matrix:
    """A sparse matrix in reference layout (1-based Int32 indices)."""
    format: int
    m: int
    n: int
    ptr: np.ndarray
    idx: np.ndarray
    val: np.ndarray
    slice_size: int = 0
    nslices: int = 0
    n_stored: int = 0
    nnz: int = 0
    _keep: list = field(default_factory=list, repr=False)

    def view(self) -> _Mat:
        for name in ("ptr", "idx"):
            a = getattr(self, name)
            assert a.dtype == np.int32 and a.flags.c_contiguous, name
        if self.val.dtype == np.bool_:
            self.val = self.val.view(np.uint8)
        mv = _Mat(self.format, _NP2CODE[self.val.dtype], self.m, self.n, self.nnz,
                  _p(self.ptr), _p(self.idx), _p(self.val), self.slice_size, self.nslices,
                  self.n_stored)
        return mv


def csr(m, n, rowptr, colval, nzval) -> Matrix:
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    colval = np.ascontiguousarray(colval, dtype=np.int32)
    nzval = np.ascontiguousarray(nzval)
    return Matrix(CSR, m, n, rowptr, colval, nzval, nnz=len(colval))


def csc(m, n, colptr, rowval, nzval) -> Matrix:
    colptr = np.ascontiguousarray(colptr, dtype=np.int32)
    rowval = np.ascontiguousarray(rowval, dtype=np.int32)
    return Matrix(CSC, m, n, colptr, rowval, np.ascontiguousarray(nzval), nnz=len(rowval))


def csr_check(m, n, rowptr, colval, nzval_len) -> int:
    """Status of the SparseGPUMatrixCSR inner-constructor checks."""
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    cv = np.ascontiguousarray(colval, dtype=np.int64)
    return lib().orc_csr_check(m, n, _p(rp), len(rp), _p(cv), len(cv), nzval_len)


def sell_from_csr(A: Matrix, slice_size: int = 32) -> Matrix:
    """Reference SELL builder, src/storage/SELL.jl:99-181."""
    L = lib()
    eff, ns, st = C.c_int64(), C.c_int64(), C.c_int64()
    _check(L.orc_sell_size(A.m, _p(A.ptr), slice_size, C.byref(eff), C.byref(ns), C.byref(st)))
    slice_ptr = np.zeros(ns.value + 1, np.int32)
    colval = np.zeros(st.value, np.int32)
    val = A.val.view(np.uint8) if A.val.dtype == np.bool_ else A.val
    nzval = np.zeros(st.value, val.dtype)
    _check(L.orc_sell_fill(A.m, _p(A.ptr), _p(A.idx), _p(val), _NP2CODE[val.dtype], eff.value,
                           ns.value, _p(slice_ptr), _p(colval), _p(nzval)))
    return Matrix(SELL, A.m, A.n, slice_ptr, colval, nzval, slice_size=eff.value,
                  nslices=ns.value, n_stored=st.value, nnz=A.nnz)


def spmv(c, A: Matrix, b, add=OP_ADD, mul=OP_MUL, accum=OP_SECOND, mask=None,
         range_=None, nthreads=0):
    """gpu_spmv! restated (in place on c, a numpy array)."""
    mv = A.view()
    b = np.ascontiguousarray(b)
    if b.dtype == np.bool_:
        b = b.view(np.uint8)
    if mask is not None:
        mask = np.ascontiguousarray(mask)
        if mask.dtype == np.bool_:
            mask = mask.view(np.uint8)
    cc = c.view(np.uint8) if c.dtype == np.bool_ else c
    r0, rl = (0, -1) if range_ is None else (range_[0], range_[1])
    _check(lib().orc_spmv(_p(cc), _NP2CODE[cc.dtype], len(cc), C.byref(mv), _p(b),
                          _NP2CODE[b.dtype], len(b), add, mul, accum, _p(mask),
                          0 if mask is None else _NP2CODE[mask.dtype],
                          0 if mask is None else len(mask), r0, rl, nthreads))
    return c


def spmm(Cm, A: Matrix, B, add=OP_ADD, mul=OP_MUL, accum=OP_SECOND, mask=None, range_=None,
         nthreads=0):
    """gpu_spmm! restated; Cm and B are Fortran-ordered (column-major) arrays."""
    mv = A.view()
    assert Cm.flags.f_contiguous and B.flags.f_contiguous
    if mask is not None:
        mask = np.ascontiguousarray(mask)
        if mask.dtype == np.bool_:
            mask = mask.view(np.uint8)
    r0, rl = (0, -1) if range_ is None else (range_[0], range_[1])
    _check(lib().orc_spmm(_p(Cm), _NP2CODE[Cm.dtype], Cm.shape[0], Cm.shape[1], C.byref(mv),
                          _p(B), _NP2CODE[B.dtype], B.shape[0], B.shape[1], add, mul, accum,
                          _p(mask), 0 if mask is None else _NP2CODE[mask.dtype],
                          0 if mask is None else len(mask), r0, rl, nthreads))
    return Cm


def bfs_distances(A: Matrix, source: int, use_mask=True, nthreads=0):
    dist = np.empty(A.m, np.int32)
    lv = C.c_int64()
    _check(lib().orc_bfs_distances(C.byref(A.view()), source, _p(dist), int(use_mask), nthreads,
                                   C.byref(lv)))
    return dist


def bfs_parents(A: Matrix, source: int, batch_size=8, nthreads=0):
    par = np.empty(A.m, np.int32)
    _check(lib().orc_bfs_parents(C.byref(A.view()), source, _p(par), batch_size, nthreads))
    return par


def dist_dtype_for(A: Matrix):
    """T_dist = ifelse(Tv == Bool, Int32, Tv), src/algorithms/shortest_path.jl:18."""
    return np.int32 if A.val.dtype in (np.bool_, np.uint8) else A.val.dtype


def shortest_path(A: Matrix, source: int, nthreads=0, max_iter=0):
    dt = np.dtype(dist_dtype_for(A))
    dist = np.empty(A.m, dt)
    it = C.c_int64()
    _check(lib().orc_shortest_path(C.byref(A.view()), source, _p(dist), _NP2CODE[dt], nthreads,
                                   max_iter, C.byref(it)))
    return dist


def shortest_path_multi(A: Matrix, sources, nthreads=0, max_iter=0):
    dt = np.dtype(dist_dtype_for(A))
    src = np.ascontiguousarray(sources, np.int32)
    dist = np.empty((A.m, len(src)), dt, order="F")
    it = C.c_int64()
    _check(lib().orc_shortest_path_multi(C.byref(A.view()), _p(src), len(src), _p(dist),
                                         _NP2CODE[dt], nthreads, max_iter, C.byref(it)))
    return dist


def spmspv(A: Matrix, b_nzind, b_nzval, b_len, add, mul, out_dtype, mask=None,
           mask_complement=False):
    """Masked SpMSpV (semantics in gg_oracle.cpp orc_spmspv); returns (nzind, nzval)."""
    mv = A.view()
    bi = np.ascontiguousarray(b_nzind, np.int32)
    bv = np.ascontiguousarray(b_nzval)
    if bv.dtype == np.bool_:
        bv = bv.view(np.uint8)
    if mask is not None:
        mask = np.ascontiguousarray(mask)
        if mask.dtype == np.bool_:
            mask = mask.view(np.uint8)
    ci = np.zeros(A.m, np.int32)
    cv = np.zeros(A.m, out_dtype)
    cnt = C.c_int64()
    _check(lib().orc_spmspv(_p(ci), _p(cv), C.byref(cnt), _NP2CODE[cv.dtype], A.m, C.byref(mv),
                            _p(bi), _p(bv), len(bi), _NP2CODE[bv.dtype], b_len, add, mul,
                            _p(mask), 0 if mask is None else _NP2CODE[mask.dtype],
                            0 if mask is None else len(mask), int(mask_complement)))
    return ci[:cnt.value].copy(), cv[:cnt.value].copy()


def bfs_levels_queue(A: Matrix, sources, nthreads=0):
    """Textbook queue BFS levels from each source (m x k, column-major Int32):
    the independent check of BFS levels at full size (orc_bfs_queue_multi)."""
    src = np.ascontiguousarray(np.atleast_1d(sources), np.int32)
    dist = np.empty((A.m, len(src)), np.int32, order="F")
    _check(lib().orc_bfs_queue_multi(C.byref(A.view()), _p(src), len(src), _p(dist), nthreads))
    return dist


def sssp_certificate(A: Matrix, sources, dist, nthreads=0):
    """Per source, the number of vertices whose distance fails the
    shortest-path certificate (0 = exact; positive weights).  dist is m x k
    (column-major) or a length-m vector for one source."""
    src = np.ascontiguousarray(np.atleast_1d(sources), np.int32)
    d = np.asfortranarray(dist.reshape(A.m, len(src), order="F"))
    bad = np.zeros(len(src), np.int64)
    _check(lib().orc_sssp_certificate(C.byref(A.view()), _p(src), len(src), _p(d),
                                      _NP2CODE[d.dtype], nthreads, _p(bad)))
    return bad


def from_scipy_csr(S, dtype=None) -> Matrix:
    """CSR Matrix (1-based) from a scipy.sparse matrix."""
    S = S.tocsr()
    S.sort_indices()
    val = S.data if dtype is None else S.data.astype(dtype)
    return csr(S.shape[0], S.shape[1], S.indptr.astype(np.int64) + 1, S.indices + 1, val)
