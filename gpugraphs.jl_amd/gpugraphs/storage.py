"""Storage types of src/storage/*.jl and src/GPUGraphsVector.jl.

Field names, index base (1) and layouts are the reference's, so a matrix's
arrays compare equal to the reference's struct fields (test/structs.jl:54-56,
204-206).  Arrays live in device memory as torch tensors on a ``ROCBackend``;
a ``CPUBackend`` keeps numpy arrays, which lets the constructors and layouts
be checked without a GPU (kernels refuse to run on it).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import ArgumentError, BoundsError, GGMatrix, GGSellHost, check, lib

_NP2CODE = {np.dtype(np.bool_): _lib.BOOL, np.dtype(np.uint8): _lib.BOOL,
            np.dtype(np.int32): _lib.INT32, np.dtype(np.int64): _lib.INT64,
            np.dtype(np.float32): _lib.FLOAT32, np.dtype(np.float64): _lib.FLOAT64}
_CODE2NP = {_lib.BOOL: np.bool_, _lib.INT32: np.int32, _lib.INT64: np.int64,
            _lib.FLOAT32: np.float32, _lib.FLOAT64: np.float64}


def dtype_code(dt) -> int:
    try:
        return _NP2CODE[np.dtype(dt)]
    except KeyError:
        raise ArgumentError(f"element type {dt} is not supported") from None


# ------------------------------------------------------------------ backends
class Backend:
    pass


class ROCBackend(Backend):
    """A ROCm device (KernelAbstractions backend counterpart)."""

    def __init__(self, device: int = 0):
        self.device = device

    def __eq__(self, o):
        return isinstance(o, ROCBackend) and o.device == self.device

    def __hash__(self):
        return hash(("roc", self.device))

    def __repr__(self):
        return f"ROCBackend({self.device})"


class CPUBackend(Backend):
    """Host arrays only: constructors and layouts work, kernels raise."""

    def __eq__(self, o):
        return isinstance(o, CPUBackend)

    def __hash__(self):
        return hash("cpu")

    def __repr__(self):
        return "CPUBackend()"


def _torch():
    import torch
    return torch


_TORCH_DT = None


def _torch_dtype(npdt):
    global _TORCH_DT
    torch = _torch()
    if _TORCH_DT is None:
        _TORCH_DT = {np.dtype(np.bool_): torch.bool, np.dtype(np.uint8): torch.uint8,
                     np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64,
                     np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64}
    return _TORCH_DT[np.dtype(npdt)]


def is_device_array(a) -> bool:
    try:
        torch = _torch()
    except ImportError:
        return False
    return isinstance(a, torch.Tensor) and a.is_cuda


def get_backend(x) -> Backend:
    if hasattr(x, "backend"):
        return x.backend
    if is_device_array(x):
        return ROCBackend(x.device.index or 0)
    return CPUBackend()


def to_backend(a, dtype, backend: Backend):
    """Array ``a`` with element type ``dtype`` on ``backend`` (copied if needed)."""
    dtype = np.dtype(dtype)
    if isinstance(backend, ROCBackend):
        torch = _torch()
        if is_device_array(a):
            t = a.to(_torch_dtype(dtype))
            return t.contiguous()
        arr = np.ascontiguousarray(np.asarray(a), dtype=dtype)
        return torch.from_numpy(arr).to(f"cuda:{backend.device}")
    if is_device_array(a):
        return np.ascontiguousarray(a.cpu().numpy(), dtype=dtype)
    return np.ascontiguousarray(np.asarray(a), dtype=dtype)


def to_host(a) -> np.ndarray:
    if is_device_array(a):
        return a.cpu().numpy()
    return np.asarray(a)


def ptr(a, host_ok: bool = False) -> int:
    """Address of an array's data (0 for None or an empty array).  Host
    arrays are accepted only for host-side checks (host_ok)."""
    if a is None:
        return 0
    if is_device_array(a):
        return a.data_ptr()
    if host_ok:
        a = np.asarray(a)
        return a.ctypes.data if a.size else 0
    raise RuntimeError("GPUGraphs kernels need device arrays (ROCBackend); got host memory")


def np_dtype_of(a):
    if is_device_array(a):
        return np.dtype(str(a.dtype).replace("torch.", ""))
    return np.asarray(a).dtype


def unit_pattern(values) -> bool:
    """True when every stored value is Bool true or an odd integer."""
    if is_device_array(values):
        torch = _torch()
        if values.numel() == 0:
            return True
        if values.dtype == torch.bool:
            return bool(values.all().item())
        if values.dtype in (torch.int32, torch.int64, torch.uint8):
            return bool(((values & 1) != 0).all().item())
        return False
    v = np.asarray(values)
    if v.size == 0:
        return True
    if v.dtype == np.bool_:
        return bool(v.all())
    if np.issubdtype(v.dtype, np.integer):
        return bool(((v & 1) != 0).all())
    return False


def _scipy_csr(S):
    S = S.tocsr(copy=True)
    S.sum_duplicates()
    S.sort_indices()
    return S


def _is_scipy(x) -> bool:
    try:
        import scipy.sparse as sp
        return sp.issparse(x)
    except ImportError:
        return False


# ------------------------------------------------------------------ matrices
class AbstractSparseGPUMatrix:
    """AbstractSparseGPUMatrix{Tv,Ti}, src/storage/GPUGraphsMatrix.jl:6-7."""
    m: int
    n: int
    backend: Backend

    def size(self, i=None):
        return (self.m, self.n) if i is None else (self.m if i == 1 else self.n)

    @property
    def shape(self):
        return (self.m, self.n)

    def __len__(self):
        return self.m * self.n

    def length(self):
        return self.m * self.n

    @property
    def dtype(self):
        return np_dtype_of(self.nzval)

    def _desc_keep(self, d: GGMatrix, keep):
        self._keep = keep
        return d

    def _check_bounds(self, i, j):
        if i < 1 or i > self.m or j < 1 or j > self.n:
            raise BoundsError(f"attempt to access {self.m}x{self.n} matrix at index [{i}, {j}]")


class SparseGPUMatrixCSR(AbstractSparseGPUMatrix):
    """SparseGPUMatrixCSR, src/storage/CSR.jl:14-136.

    Forms (backend last, as in the reference):
      SparseGPUMatrixCSR(m, n, rowptr, colval, nzval, backend)
      SparseGPUMatrixCSR(S, backend)          S: scipy sparse (its CSR; pass S.T
                                              for the reference's transpose(A))
      SparseGPUMatrixCSR(dense, backend[, Ti])
      SparseGPUMatrixCSR(Tv, Ti, backend)     empty 0 x 0
      SparseGPUMatrixCSR(m, n, Tv, Ti, backend)
    """

    def __init__(self, *args):
        if len(args) == 6:
            self._init_arrays(*args)
        elif len(args) == 5:
            m, n, Tv, Ti, backend = args
            self._init_arrays(m, n, np.ones(m + 1, np.int32), np.zeros(0, np.int32),
                              np.zeros(0, Tv), backend)
        elif len(args) == 3 and isinstance(args[2], Backend) and isinstance(args[0], type):
            Tv, Ti, backend = args
            self._init_arrays(0, 0, np.ones(1, np.int32), np.zeros(0, np.int32),
                              np.zeros(0, Tv), backend)
        elif len(args) in (2, 3):
            A, backend = args[0], args[1]
            if _is_scipy(A):
                S = _scipy_csr(A)
            else:
                import scipy.sparse as sp
                S = _scipy_csr(sp.csr_matrix(np.asarray(A)))
            self._init_arrays(S.shape[0], S.shape[1], S.indptr.astype(np.int64) + 1,
                              S.indices.astype(np.int64) + 1, S.data, backend)
        else:
            raise ArgumentError("unrecognised SparseGPUMatrixCSR constructor arguments")

    def _init_arrays(self, m, n, rowptr, colval, nzval, backend):
        # inner-constructor checks, src/storage/CSR.jl:48-63
        rp = np.ascontiguousarray(to_host(rowptr), dtype=np.int64)
        cv = np.ascontiguousarray(to_host(colval), dtype=np.int64)
        nzlen = len(nzval)
        check(lib().gg_csr_check(int(m), int(n), rp.ctypes.data, len(rp), cv.ctypes.data, len(cv),
                                 nzlen), "SparseGPUMatrixCSR")
        vdt = np_dtype_of(nzval)
        if vdt == np.uint8:
            vdt = np.dtype(np.bool_)
        self.m, self.n, self.backend = int(m), int(n), backend
        self.rowptr = to_backend(rowptr, np.int32, backend)
        self.colval = to_backend(colval, np.int32, backend)
        self.nzval = to_backend(nzval, vdt, backend)
        self.unit = unit_pattern(self.nzval)

    def nnz(self):
        return len(self.nzval)

    def descriptor(self, host_ok: bool = False) -> GGMatrix:
        def P(a):
            return ptr(a, host_ok)

        d = GGMatrix()
        d.format = _lib.CSR
        d.val_dtype = dtype_code(np_dtype_of(self.nzval))
        d.m, d.n, d.nnz = self.m, self.n, self.nnz()
        d.ptr, d.idx, d.val = P(self.rowptr), P(self.colval), P(self.nzval)
        d.flags = _lib.FLAG_UNIT_PATTERN if self.unit else 0
        if getattr(self, "symmetric", False):
            d.flags |= _lib.FLAG_SYMMETRIC
        return d

    def __getitem__(self, ij):
        i, j = ij
        self._check_bounds(i, j)
        rp, cv, nz = to_host(self.rowptr), to_host(self.colval), to_host(self.nzval)
        for p in range(rp[i - 1] - 1, rp[i] - 1):
            if cv[p] == j:
                return nz[p]
        return nz.dtype.type(0)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((to_host(self.nzval), to_host(self.colval) - 1,
                              to_host(self.rowptr).astype(np.int64) - 1), shape=(self.m, self.n))

    def __repr__(self):
        return (f"SparseGPUMatrixCSR{{{self.dtype}}}({self.m}, {self.n}) - {self.nnz()} "
                "explicit elements")


class SparseGPUMatrixCSC(AbstractSparseGPUMatrix):
    """SparseGPUMatrixCSC, src/storage/CSC.jl:2-107 (same forms as CSR)."""

    def __init__(self, *args):
        if len(args) == 6:
            self._init_arrays(*args)
        elif len(args) == 5:
            m, n, Tv, Ti, backend = args
            self._init_arrays(m, n, np.ones(n + 1, np.int32), np.zeros(0, np.int32),
                              np.zeros(0, Tv), backend)
        elif len(args) == 3 and isinstance(args[2], Backend) and isinstance(args[0], type):
            Tv, Ti, backend = args
            self._init_arrays(0, 0, np.ones(1, np.int32), np.zeros(0, np.int32),
                              np.zeros(0, Tv), backend)
        elif len(args) in (2, 3):
            import scipy.sparse as sp
            A, backend = args[0], args[1]
            S = A if _is_scipy(A) else sp.csr_matrix(np.asarray(A))
            S = S.tocsc(copy=True)
            S.sum_duplicates()
            S.sort_indices()
            self._init_arrays(S.shape[0], S.shape[1], S.indptr.astype(np.int64) + 1,
                              S.indices.astype(np.int64) + 1, S.data, backend)
        else:
            raise ArgumentError("unrecognised SparseGPUMatrixCSC constructor arguments")

    def _init_arrays(self, m, n, colptr, rowval, nzval, backend):
        # src/storage/CSC.jl:31-46 (column pointer length checked against n + 1)
        cp = np.ascontiguousarray(to_host(colptr), dtype=np.int64)
        rv = np.ascontiguousarray(to_host(rowval), dtype=np.int64)
        check(lib().gg_csr_check(int(n), int(m), cp.ctypes.data, len(cp), rv.ctypes.data, len(rv),
                                 len(nzval)), "SparseGPUMatrixCSC")
        vdt = np_dtype_of(nzval)
        self.m, self.n, self.backend = int(m), int(n), backend
        self.colptr = to_backend(colptr, np.int32, backend)
        self.rowval = to_backend(rowval, np.int32, backend)
        self.nzval = to_backend(nzval, vdt, backend)
        self.unit = unit_pattern(self.nzval)

    def nnz(self):
        return len(self.nzval)

    def descriptor(self, host_ok: bool = False) -> GGMatrix:
        def P(a):
            return ptr(a, host_ok)

        d = GGMatrix()
        d.format = _lib.CSC
        d.val_dtype = dtype_code(np_dtype_of(self.nzval))
        d.m, d.n, d.nnz = self.m, self.n, self.nnz()
        d.ptr, d.idx, d.val = P(self.colptr), P(self.rowval), P(self.nzval)
        d.flags = _lib.FLAG_UNIT_PATTERN if self.unit else 0
        if getattr(self, "symmetric", False):
            d.flags |= _lib.FLAG_SYMMETRIC
        return d

    def __getitem__(self, ij):
        i, j = ij
        self._check_bounds(i, j)
        cp, rv, nz = to_host(self.colptr), to_host(self.rowval), to_host(self.nzval)
        for p in range(cp[j - 1] - 1, cp[j] - 1):
            if rv[p] == i:
                return nz[p]
        return nz.dtype.type(0)

    def __repr__(self):
        return (f"SparseGPUMatrixCSC{{{self.dtype}}}({self.m}, {self.n}) - {self.nnz()} "
                "explicit elements")


class SparseGPUMatrixSELL(AbstractSparseGPUMatrix):
    """SparseGPUMatrixSELL, src/storage/SELL.jl:3-247.

      SparseGPUMatrixSELL(S, backend, slice_size=32, sigma=1, heavy_threshold=0)
      SparseGPUMatrixSELL(dense, backend, slice_size=32, ...)
      SparseGPUMatrixSELL(A_csr, ..)   from a SparseGPUMatrixCSR
      SparseGPUMatrixSELL(m, n, perm, slice_size, nslices, nnz, n_stored,
                          slice_ptr, colval, nzval, backend)   raw arrays

    sigma = 1 with no heavy threshold is the reference layout (identity perm,
    slice_size clipped to the row count, src/storage/SELL.jl:108,115).
    sigma > 1 sorts rows by length inside windows of sigma rows; rows longer
    than heavy_threshold are held in a CSR partition processed in segments of
    heavy_seg_len entries.  relabel=True (square matrices) sorts all rows by
    length and stores columns as internal ids (slot index, then heavy index),
    so a row's neighbours are gathered from the compact, degree-ordered state
    arrays; every operation still takes and returns vectors in original order.
    sort_entries=True stores each row's entries by increasing stored column
    (when relabelled: light rows by decreasing degree, then the heavy rows): folds visit a row's terms in that order.
    rcm=True (with relabel) lays out rows of equal length in reverse
    Cuthill-McKee order instead of row order (host builder only).
    """

    def __init__(self, *args, slice_size=None, sigma=1, heavy_threshold=0, heavy_seg_len=4096,
                 relabel=False, sort_entries=False, rcm=False):
        if len(args) == 11:
            self._init_raw(*args)
            return
        if len(args) < 2:
            raise ArgumentError("unrecognised SparseGPUMatrixSELL constructor arguments")
        A, backend = args[0], args[1]
        if len(args) >= 3:
            slice_size = args[2]
        if slice_size is None:
            slice_size = 32
        if isinstance(A, SparseGPUMatrixCSR):
            m, n = A.m, A.n
            rowptr = np.ascontiguousarray(to_host(A.rowptr), np.int32)
            colval = np.ascontiguousarray(to_host(A.colval), np.int32)
            nzval = np.ascontiguousarray(to_host(A.nzval))
        else:
            import scipy.sparse as sp
            S = _scipy_csr(A if _is_scipy(A) else sp.csr_matrix(np.asarray(A)))
            m, n = S.shape
            rowptr = (S.indptr.astype(np.int64) + 1).astype(np.int32)
            colval = (S.indices + 1).astype(np.int32)
            nzval = np.ascontiguousarray(S.data)
        self._build(m, n, rowptr, colval, nzval, backend, int(slice_size), int(sigma),
                    int(heavy_threshold), int(heavy_seg_len), bool(relabel), bool(sort_entries),
                    bool(rcm))

    def _build(self, m, n, rowptr, colval, nzval, backend, slice_size, sigma, heavy_threshold,
               heavy_seg_len, relabel=False, sort_entries=False, rcm=False):
        vdt = nzval.dtype
        if vdt == np.bool_:
            nzval = nzval.view(np.uint8)
        out = C.POINTER(GGSellHost)()
        check(lib().gg_sell_build(m, n, rowptr.ctypes.data, colval.ctypes.data, nzval.ctypes.data,
                                  dtype_code(vdt), slice_size, sigma, heavy_threshold,
                                  heavy_seg_len,
                                  (_lib.SELL_RELABEL if relabel else 0)
                                  | (_lib.SELL_SORT_ENTRIES if sort_entries else 0)
                                  | (_lib.SELL_RCM if rcm else 0),
                                  C.byref(out)), "SparseGPUMatrixSELL")
        try:
            h = out.contents

            def arr(p, count, dt):
                if not p or count == 0:
                    return np.zeros(0, dt)
                nbytes = count * np.dtype(dt).itemsize
                return np.frombuffer((C.c_char * nbytes).from_address(p), dtype=dt).copy()

            self.m, self.n, self.backend = int(m), int(n), backend
            self.slice_size, self.nslices = int(h.slice_size), int(h.nslices)
            self._nnz, self.n_stored, self.sigma = int(h.nnz), int(h.n_stored), sigma
            self.slice_ptr = to_backend(arr(h.slice_ptr, h.nslices + 1, np.int32), np.int32, backend)
            self.colval = to_backend(arr(h.colval, h.n_stored, np.int32), np.int32, backend)
            self.nzval = to_backend(arr(h.nzval, h.n_stored, vdt), vdt, backend)
            if h.perm:
                self.perm = arr(h.perm, h.n_slots, np.int32)
                self.slot_len = to_backend(arr(h.slot_len, h.n_slots, np.int32), np.int32, backend)
            else:
                self.perm = np.arange(1, m + 1, dtype=np.int64)   # identity, src/storage/SELL.jl:115
                self.slot_len = None
            self._perm_dev = (to_backend(self.perm, np.int32, backend) if h.perm else None)
            self.n_heavy, self.heavy_seg_len, self.n_seg = int(h.n_heavy), int(h.heavy_seg_len), int(h.n_seg)
            if h.n_heavy:
                self.heavy_row = to_backend(arr(h.heavy_row, h.n_heavy, np.int32), np.int32, backend)
                self.heavy_ptr = to_backend(arr(h.heavy_ptr, h.n_heavy + 1, np.int64), np.int64, backend)
                self.heavy_seg = to_backend(arr(h.heavy_seg, h.n_heavy + 1, np.int32), np.int32, backend)
                self.heavy_col = to_backend(arr(h.heavy_col, h.heavy_nnz, np.int32), np.int32, backend)
                self.heavy_val = to_backend(arr(h.heavy_val, h.heavy_nnz, vdt), vdt, backend)
            else:
                self.heavy_row = self.heavy_ptr = self.heavy_seg = None
                self.heavy_col = self.heavy_val = None
            if h.vinv:   # original row -> internal id (1-based), host and device copies
                self.vinv = arr(h.vinv, m, np.int32)
                self._vinv_dev = to_backend(self.vinv, np.int32, backend)
            else:
                self.vinv = self._vinv_dev = None
        finally:
            lib().gg_sell_host_free(out)
        self.unit = unit_pattern(nzval.view(np.bool_) if vdt == np.bool_ else nzval)

    @classmethod
    def build_on_device(cls, A, slice_size=32, sigma=1, heavy_threshold=0, heavy_seg_len=4096,
                        relabel=False, sort_entries=False, rcm=False):
        """The same layouts as the constructor, built in device memory from a
        SparseGPUMatrixCSR on a ROCBackend (gg_sell_plan_device /
        gg_sell_fill_device): the rows never visit the host."""
        if not isinstance(A, SparseGPUMatrixCSR) or not isinstance(A.backend, ROCBackend):
            raise ArgumentError("build_on_device takes a SparseGPUMatrixCSR on a ROCBackend")
        torch = _torch()
        backend = A.backend
        ctx = _lib.context(backend.device)
        vdt = np_dtype_of(A.nzval)
        plan = C.c_void_p()
        z = _lib.GGSellSizes()
        check(lib().gg_sell_plan_device(ctx, A.m, A.n, ptr(A.rowptr), ptr(A.colval),
                                        ptr(A.nzval), dtype_code(vdt), int(slice_size),
                                        int(sigma), int(heavy_threshold), int(heavy_seg_len),
                                        (_lib.SELL_RELABEL if relabel else 0)
                                        | (_lib.SELL_SORT_ENTRIES if sort_entries else 0)
                                        | (_lib.SELL_RCM if rcm else 0),
                                        C.byref(plan),
                                        C.byref(z)), "SparseGPUMatrixSELL")
        try:
            dev = torch.device("cuda", backend.device)

            def empty(count, dt):
                return torch.empty(max(int(count), 1), dtype=_torch_dtype(dt), device=dev)[:int(count)]

            M = cls.__new__(cls)
            M.m, M.n, M.backend = int(A.m), int(A.n), backend
            M.slice_size, M.nslices, M.n_stored = int(z.slice_size), int(z.nslices), int(z.n_stored)
            M.sigma = int(sigma)
            M.slice_ptr = empty(z.nslices + 1, np.int32)
            M.colval = empty(z.n_stored, np.int32)
            M.nzval = empty(z.n_stored, vdt)
            perm_dev = empty(z.n_slots, np.int32) if z.has_perm else None
            M.slot_len = empty(z.n_slots, np.int32) if z.has_perm else None
            M.n_heavy, M.n_seg = int(z.n_heavy), int(z.n_seg)
            M.heavy_seg_len = int(heavy_seg_len)
            if z.n_heavy:
                M.heavy_row = empty(z.n_heavy, np.int32)
                M.heavy_ptr = empty(z.n_heavy + 1, np.int64)
                M.heavy_seg = empty(z.n_heavy + 1, np.int32)
                M.heavy_col = empty(z.heavy_nnz, np.int32)
                M.heavy_val = empty(z.heavy_nnz, vdt)
            else:
                M.heavy_row = M.heavy_ptr = M.heavy_seg = M.heavy_col = M.heavy_val = None
            vinv_dev = empty(A.m, np.int32) if z.has_vinv else None
            d = GGMatrix()
            d.ptr, d.idx, d.val = ptr(M.slice_ptr), ptr(M.colval), ptr(M.nzval)
            if perm_dev is not None:
                d.perm, d.slot_len = ptr(perm_dev), ptr(M.slot_len)
            if z.n_heavy:
                d.heavy_row, d.heavy_ptr, d.heavy_seg = (ptr(M.heavy_row), ptr(M.heavy_ptr),
                                                         ptr(M.heavy_seg))
                d.heavy_col, d.heavy_val = ptr(M.heavy_col), ptr(M.heavy_val)
            if vinv_dev is not None:
                d.vinv = ptr(vinv_dev)
            check(lib().gg_sell_fill_device(plan, C.byref(d)), "SparseGPUMatrixSELL")
        finally:
            lib().gg_sell_plan_free(plan)
        M._nnz = int(d.nnz)
        M._perm_dev = perm_dev
        M.perm = perm_dev.cpu().numpy() if perm_dev is not None else np.arange(1, A.m + 1,
                                                                               dtype=np.int64)
        M._vinv_dev = vinv_dev
        M.vinv = vinv_dev.cpu().numpy() if vinv_dev is not None else None
        M.unit = A.unit
        M.symmetric = getattr(A, "symmetric", False)
        return M

    def _init_raw(self, m, n, perm, slice_size, nslices, nnz, n_stored, slice_ptr, colval, nzval,
                  backend):
        # inner-constructor checks, src/storage/SELL.jl:42-52
        cv = np.ascontiguousarray(to_host(colval), dtype=np.int64)
        check(lib().gg_sell_check(int(n), int(nslices), int(n_stored), len(slice_ptr),
                                  cv.ctypes.data, len(cv), len(nzval)), "SparseGPUMatrixSELL")
        vdt = np_dtype_of(nzval)
        # beyond the reference's checks: the kernels index through slice_ptr
        spt = np.asarray(to_host(slice_ptr), dtype=np.int64)
        if (len(spt) and (spt[0] != 1 or spt[-1] != int(n_stored) + 1 or np.any(np.diff(spt) < 0)
                          or np.any(np.diff(spt) % max(int(slice_size), 1)))) \
                or int(nslices) * int(slice_size) < int(m):
            raise ArgumentError("slice_ptr is not a SELL layout of these sizes")
        self.m, self.n, self.backend = int(m), int(n), backend
        self.perm = np.asarray(perm)
        if not np.array_equal(self.perm, np.arange(1, m + 1)):
            raise ArgumentError("a raw SELL matrix must use the identity permutation")
        self.slice_size, self.nslices, self._nnz = int(slice_size), int(nslices), int(nnz)
        self.n_stored, self.sigma = int(n_stored), 1
        self.slice_ptr = to_backend(slice_ptr, np.int32, backend)
        self.colval = to_backend(colval, np.int32, backend)
        self.nzval = to_backend(nzval, vdt, backend)
        self.slot_len = self._perm_dev = None
        self.vinv = self._vinv_dev = None
        self.n_heavy, self.heavy_seg_len, self.n_seg = 0, 0, 0
        self.heavy_row = self.heavy_ptr = self.heavy_seg = self.heavy_col = self.heavy_val = None
        h = to_host(self.nzval)
        self.unit = unit_pattern(h[to_host(self.colval) != -1])

    def nnz(self):
        return self._nnz   # true nonzeros, src/storage/SELL.jl:242

    def descriptor(self, host_ok: bool = False) -> GGMatrix:
        def P(a):
            return ptr(a, host_ok)

        d = GGMatrix()
        d.format = _lib.SELL
        d.val_dtype = dtype_code(np_dtype_of(self.nzval))
        d.m, d.n, d.nnz = self.m, self.n, self._nnz
        d.ptr, d.idx, d.val = P(self.slice_ptr), P(self.colval), P(self.nzval)
        d.slice_size, d.nslices, d.n_stored = self.slice_size, self.nslices, self.n_stored
        d.perm = P(self._perm_dev) if self._perm_dev is not None else 0
        d.slot_len = P(self.slot_len) if self.slot_len is not None else 0
        d.n_heavy = self.n_heavy
        if self.n_heavy:
            d.heavy_row, d.heavy_ptr = P(self.heavy_row), P(self.heavy_ptr)
            d.heavy_seg, d.heavy_seg_len = P(self.heavy_seg), self.heavy_seg_len
            d.heavy_nseg = self.n_seg
            d.heavy_col, d.heavy_val = P(self.heavy_col), P(self.heavy_val)
        d.flags = _lib.FLAG_UNIT_PATTERN if self.unit else 0
        if getattr(self, "symmetric", False):
            d.flags |= _lib.FLAG_SYMMETRIC
        if self._vinv_dev is not None:
            d.flags |= _lib.FLAG_INTERNAL_IDS
            d.vinv = P(self._vinv_dev)
        return d

    def __getitem__(self, ij):
        i, j = ij
        self._check_bounds(i, j)
        if self.vinv is not None:
            j = int(self.vinv[j - 1])   # columns are stored as internal ids
        sp_ = to_host(self.slice_ptr)
        cv, nz = to_host(self.colval), to_host(self.nzval)
        if self.n_heavy:
            hr = to_host(self.heavy_row)
            hit = np.nonzero(hr == i)[0]
            if len(hit):
                hp = to_host(self.heavy_ptr)
                hc, hv = to_host(self.heavy_col), to_host(self.heavy_val)
                h = hit[0]
                for p in range(hp[h], hp[h + 1]):
                    if hc[p] == j:
                        return hv[p]
                return hv.dtype.type(0)
        slots = np.nonzero(np.asarray(self.perm) == i)[0]
        if len(slots) == 0:
            return nz.dtype.type(0)
        slot = slots[0]
        C_ = self.slice_size
        s, r = divmod(int(slot), C_)
        for p in range(sp_[s] - 1 + r, sp_[s + 1] - 1, C_):
            if cv[p] == -1:
                break
            if cv[p] == j:
                return nz[p]
        return nz.dtype.type(0)

    def __repr__(self):
        return (f"SparseGPUMatrixSELL{{{self.dtype} - int32}}({self.m}, {self.n}) - "
                f"{self._nnz} explicit elements")


def sprand_gpu(Tv, m: int, n: int, p: float, backend: Backend, seed=None):
    """CSR of transpose(sprand(Tv, m, n, p)), src/storage/CSR.jl:185-186."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    S = sp.random(m, n, density=p, format="csc", random_state=rng, dtype=np.float64)
    if np.dtype(Tv) == np.bool_:
        S.data[:] = 1
    else:
        S.data = rng.random(S.nnz)
    return SparseGPUMatrixCSR(S.T.astype(Tv), backend)


# ------------------------------------------------------------------- vector
class SparseGPUVector:
    """SparseGPUVector, src/GPUGraphsVector.jl:17-129.

      SparseGPUVector(n, nzind, nzval, backend)
      SparseGPUVector(dense_vector, backend)     stores the nonzeros
      SparseGPUVector(Tv, Ti, backend)           empty
    """

    def __init__(self, *args):
        if len(args) == 4:
            n, nzind, nzval, backend = args
        elif len(args) == 3 and isinstance(args[2], Backend):
            Tv, Ti, backend = args
            n, nzind, nzval = 0, np.zeros(0, np.int32), np.zeros(0, Tv)
        elif len(args) == 2:
            v, backend = args
            if _is_scipy(v):
                v = v.tocsr() if v.shape[0] == 1 else v.T.tocsr()
                v.sort_indices()
                n, nzind, nzval = v.shape[1], v.indices + 1, v.data
            else:
                v = to_host(v)
                nzind = np.nonzero(v)[0] + 1
                n, nzval = len(v), v[nzind - 1]
        else:
            raise ArgumentError("unrecognised SparseGPUVector constructor arguments")
        if len(nzind) != len(nzval):
            raise ArgumentError("length(nzind) must be equal to length(nzval)")
        self.n, self.backend = int(n), backend
        self.nzind = to_backend(nzind, np.int32, backend)
        self.nzval = to_backend(nzval, np_dtype_of(nzval), backend)

    def size(self, i=None):
        return (self.n,) if i is None else (self.n if i == 1 else 1)

    def __len__(self):
        return self.n

    def length(self):
        return self.n

    def nnz(self):
        return len(self.nzval)

    def __getitem__(self, i):
        if i < 1 or i > self.n:
            raise BoundsError(f"attempt to access {self.n}-element vector at index [{i}]")
        ind = to_host(self.nzind)
        pos = np.nonzero(ind == i)[0]
        vals = to_host(self.nzval)
        return vals[pos[0]] if len(pos) else vals.dtype.type(0)

    def __setitem__(self, i, val):
        if i < 1 or i > self.n:
            raise BoundsError(f"attempt to access {self.n}-element vector at index [{i}]")
        pos = np.nonzero(to_host(self.nzind) == i)[0]
        if len(pos) == 0:
            raise ArgumentError(f"Index ({i}) is not in the vector. Adding new values is not "
                                "supported yet.")
        self.nzval[int(pos[0])] = val

    def __repr__(self):
        return f"SparseGPUVector{{{np_dtype_of(self.nzval)}}} ({self.n},) - {self.nnz()} explicit elements"
