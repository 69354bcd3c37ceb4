"""gpu_spmv! / gpu_spmm! / gpu_spmspv! (src/spmv, src/spmm, src/spmspv.jl).

Julia's ``!`` suffix becomes a trailing underscore: ``gpu_spmv_(C, A, B; ...)``
mutates ``C`` in place exactly as ``gpu_spmv!`` does.  A ``range_`` is a
Python ``range`` of 1-based rows (``range(2, 6)`` is Julia's ``2:5``).
Every call is enqueued on the torch current stream of the matrix's device and
returns without waiting, like the reference's asynchronous launch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import ArgumentError, DimensionMismatch, check, context, lib
from .algebra import GPUGraphs_add, GPUGraphs_mul, GPUGraphs_second, op_code
from .storage import (ROCBackend, SparseGPUMatrixCSC, SparseGPUVector, dtype_code, get_backend,
                      np_dtype_of, ptr)


def _same_backend(*xs):
    bs = {get_backend(x) for x in xs if x is not None}
    if len(bs) != 1:
        raise ArgumentError("All inputs must be on the same backend")
    b = bs.pop()
    if not isinstance(b, ROCBackend):
        raise RuntimeError("GPUGraphs kernels need a ROCBackend (device arrays); got " + repr(b))
    return b


def _range_args(r):
    if r is None:
        return 0, -1
    if isinstance(r, range):
        if r.step != 1:
            raise ArgumentError("range must be contiguous")
        return r.start, len(r)
    start, stop = r  # inclusive (start, stop) pair
    return start, stop - start + 1


def gpu_spmv_(C_, A, B, mul=GPUGraphs_mul, add=GPUGraphs_add, accum=GPUGraphs_second,
              range_=None, mask=None):
    """C = accum(C, A add.mul B) on the rows of mask / range_ (src/spmv/CSR.jl:2-29).

    A SparseGPUVector mask of length m lists the rows to compute (its stored
    indices, distinct): the sparse-mask CSR kernel of src/spmv/CSR.jl:90-115."""
    if isinstance(mask, SparseGPUVector):
        if range_ is not None:
            raise ArgumentError("a sparse mask takes no range")
        if mask.n != A.m:
            raise DimensionMismatch("Mask length must be equal to the number of rows in A")
        b = _same_backend(A, B, C_)
        d = A.descriptor()
        check(lib().gg_spmv_rows(context(b.device), ptr(C_), dtype_code(np_dtype_of(C_)),
                                 C_.numel(), C.byref(d), ptr(B), dtype_code(np_dtype_of(B)),
                                 B.numel(), op_code(add), op_code(mul), op_code(accum),
                                 ptr(mask.nzind), mask.nnz()), "gpu_spmv!")
        return None
    b = _same_backend(A, B, C_, mask)
    d = A.descriptor()
    r0, rl = _range_args(range_)
    # a CSC matrix ignores add and accum, as src/spmv/CSC.jl:14-20 does
    ctx = context(b.device)
    check(lib().gg_spmv(ctx, ptr(C_), dtype_code(np_dtype_of(C_)), C_.numel(), C.byref(d), ptr(B),
                        dtype_code(np_dtype_of(B)), B.numel(), op_code(add), op_code(mul),
                        op_code(accum), ptr(mask) if mask is not None else None,
                        dtype_code(np_dtype_of(mask)) if mask is not None else 0,
                        mask.numel() if mask is not None else 0, r0, rl), "gpu_spmv!")
    return None


def _layout(t):
    if t.dim() != 2:
        raise DimensionMismatch("SpMM operands are matrices")
    if t.is_contiguous():
        return _lib.ROW_MAJOR
    if t.T.is_contiguous():
        return _lib.COL_MAJOR
    raise ArgumentError("SpMM operands must be row-major or column-major contiguous")


def gpu_spmm_(C_, A, B, mul=GPUGraphs_mul, add=GPUGraphs_add, accum=GPUGraphs_second,
              mask=None, range_=None):
    """C = accum(C, A add.mul B) for a dense B (src/spmm/CSR.jl:111-204).

    Julia matrices are column-major: a torch tensor ``t.T.contiguous().T``
    (Fortran order) has the reference's layout; a contiguous tensor is
    row-major.  Both are accepted for C and B.
    """
    b = _same_backend(A, B, C_, mask)
    d = A.descriptor()
    r0, rl = _range_args(range_)
    ctx = context(b.device)
    check(lib().gg_spmm(ctx, ptr(C_), dtype_code(np_dtype_of(C_)), C_.shape[0], C_.shape[1],
                        _layout(C_), C.byref(d), ptr(B), dtype_code(np_dtype_of(B)), B.shape[0],
                        B.shape[1], _layout(B), op_code(add), op_code(mul), op_code(accum),
                        ptr(mask) if mask is not None else None,
                        dtype_code(np_dtype_of(mask)) if mask is not None else 0,
                        mask.numel() if mask is not None else 0, r0, rl), "gpu_spmm!")
    return None


def gpu_spmspv(A, B: SparseGPUVector, mul=GPUGraphs_mul, add=GPUGraphs_add, out_dtype=None,
               mask=None, mask_complement=False) -> SparseGPUVector:
    """Masked SpMSpV (the product src/spmspv.jl:23-54 intends).

    Returns a new SparseGPUVector of the rows that received a contribution
    (indices ascending).  A must be a SparseGPUMatrixCSC (for a CSR A_T of a
    symmetric graph, ``csc_view`` gives the CSC view without copying).
    """
    import torch
    b = _same_backend(A, B.nzind, mask)
    if not isinstance(A, (SparseGPUMatrixCSC, CSCView)):
        raise ArgumentError("gpu_spmspv expands the columns of a CSC matrix")
    d = A.descriptor()
    vt = np_dtype_of(A.nzval)
    bt = np_dtype_of(B.nzval)
    if out_dtype is None:
        out_dtype = np.int32 if mul.code == 14 else np.result_type(
            np.bool_ if vt == np.bool_ else vt, bt)
    dev = f"cuda:{b.device}"
    from .storage import _torch_dtype
    ci = torch.empty(max(A.m, 1), dtype=torch.int32, device=dev)
    cv = torch.empty(max(A.m, 1), dtype=_torch_dtype(out_dtype), device=dev)
    cnt = C.c_int64()
    ctx = context(b.device)
    check(lib().gg_spmspv(ctx, ptr(ci), ptr(cv), C.byref(cnt), dtype_code(out_dtype), A.m,
                          C.byref(d), ptr(B.nzind), ptr(B.nzval), B.nnz(), dtype_code(bt), B.n,
                          op_code(add), op_code(mul), ptr(mask) if mask is not None else None,
                          dtype_code(np_dtype_of(mask)) if mask is not None else 0,
                          mask.numel() if mask is not None else 0, int(mask_complement)),
          "gpu_spmspv")
    k = cnt.value
    return SparseGPUVector(A.m, ci[:k].clone(), cv[:k].clone(), b)


gpu_spmspv_ = gpu_spmspv


class CSCView:
    """The CSC of an m x n matrix given by the CSR arrays of its transpose
    (zero copy): for a symmetric A_T, csc_view(A_T) is A_T column-wise."""

    def __init__(self, csr_of_transpose):
        t = csr_of_transpose
        self.m, self.n = t.n, t.m
        self.colptr, self.rowval, self.nzval = t.rowptr, t.colval, t.nzval
        self.backend = t.backend
        self.unit = t.unit

    def nnz(self):
        return len(self.nzval)

    def descriptor(self):
        d = _lib.GGMatrix()
        d.format = _lib.CSC
        d.val_dtype = dtype_code(np_dtype_of(self.nzval))
        d.m, d.n, d.nnz = self.m, self.n, self.nnz()
        d.ptr, d.idx, d.val = ptr(self.colptr), ptr(self.rowval), ptr(self.nzval)
        d.flags = _lib.FLAG_UNIT_PATTERN if self.unit else 0
        return d


def csc_view(csr_of_transpose) -> CSCView:
    return CSCView(csr_of_transpose)
