"""CPU tests of the product's host side: the C-ABI library loads and exports
every symbol of include/gpugraphs.h, the storage constructors reproduce the
reference's layouts and errors, the SELL-C-sigma hybrid keeps every row, and
the generators produce the graphs they promise.  No kernel runs here."""
import json
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import GOLDEN, ROOT

import gpugraphs as gg
from gpugraphs import _lib
from gpugraphs import generators as G

CPU = gg.CPUBackend()


def test_library_exports_every_header_symbol():
    hdr = open(os.path.join(ROOT, "include", "gpugraphs.h")).read()
    names = set(re.findall(r"\b(gg_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) > 25
    L = gg.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.SIGNATURES) == names, set(_lib.SIGNATURES) ^ names
    assert L.gg_abi_version() == 1


def test_status_strings_name_julia_exceptions():
    L = gg.lib()
    assert L.gg_status_string(1) == b"DimensionMismatch"
    assert L.gg_status_string(2) == b"ArgumentError"
    assert L.gg_status_string(3) == b"AssertionError"
    assert L.gg_status_string(4) == b"BoundsError"


def test_sell_known_answer_through_product_builder():
    """test/structs.jl:176-190 through gg_sell_build, all three constructor routes."""
    kat = json.load(open(os.path.join(GOLDEN, "sell_kat.json")))
    dense = np.array(kat["dense"], np.float32)
    e = kat["expected"]
    routes = [gg.SparseGPUMatrixSELL(sp.csc_matrix(dense), CPU),
              gg.SparseGPUMatrixSELL(sp.csc_matrix(dense.T).T, CPU),
              gg.SparseGPUMatrixSELL(dense, CPU)]
    for A in routes:
        assert A.size() == (4, 4) and A.length() == 16 and A.nnz() == e["nnz"]
        assert A.slice_size == e["slice_size"] and A.n_stored == e["n_stored"]
        assert A.slice_ptr.tolist() == e["slice_ptr"]
        assert A.colval.tolist() == e["colval"]
        assert A.nzval.tolist() == e["nzval"]
        for i in range(1, 5):
            for j in range(1, 5):
                assert A[i, j] == dense[i - 1, j - 1]
        with pytest.raises(gg.BoundsError):
            A[11, 1]


def test_csr_constructor_errors_through_product():
    """test/structs.jl:79-150 through gg_csr_check."""
    cases = json.load(open(os.path.join(GOLDEN, "csr_errors.json")))["cases"]
    for c in cases:
        nz = np.arange(c["nzval_len"], dtype=np.float32)
        if c["expect"] == "ArgumentError":
            with pytest.raises(gg.ArgumentError):
                gg.SparseGPUMatrixCSR(c["m"], c["n"], c["rowptr"], c["colval"], nz, CPU)
        else:
            A = gg.SparseGPUMatrixCSR(c["m"], c["n"], c["rowptr"], c["colval"], nz, CPU)
            assert A.size() == (c["m"], c["n"])


def test_csr_constructor_routes_match_fields():
    """test/structs.jl:30-78: the four routes give the same fields."""
    rng = np.random.default_rng(0)
    Acsc = sp.random(10, 10, density=0.5, format="csc", random_state=rng, dtype=np.float32)
    Acsr_t = Acsc.T.tocsc()  # sparse(transpose(A_csc))
    ref_rowptr = Acsr_t.indptr + 1
    ref_colval = Acsr_t.indices + 1
    ref_nzval = Acsr_t.data
    B0 = gg.SparseGPUMatrixCSR(10, 10, ref_rowptr, ref_colval, ref_nzval, CPU)
    B1 = gg.SparseGPUMatrixCSR(Acsc, CPU)
    B2 = gg.SparseGPUMatrixCSR(Acsr_t.T, CPU)
    B3 = gg.SparseGPUMatrixCSR(Acsc.toarray(), CPU)
    for B in (B0, B1, B2, B3):
        assert B.size() == (10, 10) and B.length() == 100 and B.nnz() == Acsc.nnz
        assert B.rowptr.tolist() == ref_rowptr.tolist()
        assert B.colval.tolist() == ref_colval.tolist()
        np.testing.assert_array_equal(B.nzval, ref_nzval)
        D = Acsc.toarray()
        assert all(B[i, j] == D[i - 1, j - 1] for i in range(1, 11) for j in range(1, 11))
        with pytest.raises(gg.BoundsError):
            B[11, 1]
    E = gg.SparseGPUMatrixCSR(np.float32, np.int32, CPU)
    assert E.size() == (0, 0) and E.nnz() == 0 and E.length() == 0
    R = gg.sprand_gpu(np.float32, 10, 10, 0.5, CPU, seed=1)
    assert R.size() == (10, 10)


def test_csc_constructor_and_vector():
    rng = np.random.default_rng(1)
    Acsc = sp.random(10, 10, density=0.5, format="csc", random_state=rng, dtype=np.float32)
    A = gg.SparseGPUMatrixCSC(Acsc, CPU)
    assert A.colptr.tolist() == (Acsc.indptr + 1).tolist()
    assert A.rowval.tolist() == (Acsc.indices + 1).tolist()
    D = Acsc.toarray()
    assert all(A[i, j] == D[i - 1, j - 1] for i in range(1, 11) for j in range(1, 11))
    v = sp.random(1, 10, density=0.5, format="csr", random_state=rng, dtype=np.float32)
    for V in (gg.SparseGPUVector(10, v.indices + 1, v.data, CPU), gg.SparseGPUVector(v, CPU),
              gg.SparseGPUVector(v.toarray()[0], CPU)):
        assert V.size() == (10,) and V.length() == 10
        assert V.nzind.tolist() == (v.indices + 1).tolist()
        np.testing.assert_array_equal(V.nzval, v.data)
    assert gg.SparseGPUVector(np.float32, np.int32, CPU).length() == 0
    with pytest.raises(gg.ArgumentError):
        gg.SparseGPUVector(10, [1, 2], [1.0], CPU)
    V = gg.SparseGPUVector(10, [2, 5], np.array([1.0, 2.0], np.float32), CPU)
    with pytest.raises(gg.ArgumentError):
        V[3] = 1.0  # adding entries is not supported (src/GPUGraphsVector.jl:113-119)


def _rows_of_sell(A):
    """Rebuild {row: sorted (col, val)} from a SELL hybrid's host arrays."""
    rows = {}
    sp_, cv, nz = A.slice_ptr, A.colval, A.nzval
    perm = np.asarray(A.perm)
    slen = None if A.slot_len is None else np.asarray(A.slot_len)
    C = A.slice_size
    for slot, row in enumerate(perm):
        if row == 0:
            continue
        s, r = divmod(slot, C)
        width = (sp_[s + 1] - sp_[s]) // C
        ent = []
        for k in range(width if slen is None else slen[slot]):
            p = sp_[s] - 1 + k * C + r
            if cv[p] == -1:
                break
            ent.append((cv[p], nz[p]))
        rows[int(row)] = ent
    if A.n_heavy:
        for h, row in enumerate(A.heavy_row):
            a, b = A.heavy_ptr[h], A.heavy_ptr[h + 1]
            rows[int(row)] = list(zip(A.heavy_col[a:b], A.heavy_val[a:b]))
    return rows


@pytest.mark.parametrize("sigma,heavy", [(1, 0), (64, 0), (4096, 0), (4096, 100), (7, 30)])
def test_sell_hybrid_keeps_every_row(sigma, heavy):
    R = G.rmat(12, seed=5, dtype=np.float32, values=G.UNIFORM01)
    A = R.to_sell(CPU, 64, sigma, heavy, heavy_seg_len=64)
    rows = _rows_of_sell(A)
    S = R.to_scipy()
    assert len(rows) == R.m
    for v in range(R.m):
        a, b = S.indptr[v], S.indptr[v + 1]
        exp = list(zip(S.indices[a:b] + 1, S.data[a:b]))
        assert rows[v + 1] == exp
    if sigma > 1:
        # inside a window, slots are ordered by decreasing length
        lens = np.asarray(A.slot_len)
        for w0 in range(0, len(lens), sigma):
            seg = [l for l, p in zip(lens[w0:w0 + sigma], A.perm[w0:w0 + sigma]) if p]
            assert seg == sorted(seg, reverse=True)
    if heavy:
        assert all(len(rows[int(r)]) > heavy for r in A.heavy_row)
        assert A.n_seg == sum(-(-len(rows[int(r)]) // 64) for r in A.heavy_row)


@pytest.mark.parametrize("heavy", [0, 40, 1])
def test_sell_relabel_keeps_every_row(heavy):
    """Degree relabelling: vinv inverts (perm, heavy_row) into internal ids
    (slot index, then slots + heavy index), slot lengths fall over the whole
    matrix, and mapping the stored internal columns back gives every row."""
    R = G.rmat(12, seed=6, dtype=np.float32, values=G.UNIFORM01)
    A = R.to_sell(CPU, 64, 1, heavy, heavy_seg_len=64, relabel=True)
    base = A.nslices * A.slice_size
    assert A.vinv is not None and len(A.vinv) == R.m
    internal_to_orig = np.zeros(base + A.n_heavy + 1, np.int64)
    for slot, p in enumerate(A.perm):
        if p:
            assert A.vinv[p - 1] == slot + 1
            internal_to_orig[slot + 1] = p
    for h, r in enumerate(A.heavy_row if A.n_heavy else []):
        assert A.vinv[r - 1] == base + h + 1
        internal_to_orig[base + h + 1] = r
    assert sorted(internal_to_orig[internal_to_orig > 0]) == list(range(1, R.m + 1))
    lens = [l for l, p in zip(A.slot_len, A.perm) if p]
    assert lens == sorted(lens, reverse=True)
    rows = _rows_of_sell(A)
    S = R.to_scipy()
    for v in range(R.m):
        a, b = S.indptr[v], S.indptr[v + 1]
        back = [(int(internal_to_orig[c]), x) for c, x in rows[v + 1]]   # input order kept
        assert back == list(zip(S.indices[a:b] + 1, S.data[a:b]))
    d = A.descriptor(host_ok=True)
    assert d.flags & 2 and d.vinv
    for i, j in [(1, 1), (5, int(S.indices[S.indptr[4]]) + 1), (R.m, R.m)]:
        assert A[i, j] == S[i - 1, j - 1]
    _walk_sell_indexing(A)


def _bandwidth(S, order):
    """Largest |pos(i) - pos(j)| over the entries when vertex order[k] (1-based) sits at k."""
    pos = np.empty(S.shape[0], np.int64)
    pos[np.asarray(order) - 1] = np.arange(S.shape[0])
    c = S.tocoo()
    return int(np.abs(pos[c.row] - pos[c.col]).max()) if c.nnz else 0


def _host(S):
    S = sp.csr_matrix(S)
    S.sort_indices()
    return G.HostCSR(S.shape[0], S.shape[1], (S.indptr + 1).astype(np.int32),
                     (S.indices + 1).astype(np.int32), S.data, True)


def test_rcm_order_bandwidth():
    """gg_rcm_order is a permutation; it lays a shuffled path out with bandwidth
    1 and a shuffled 20 x 30 grid with bandwidth at most 2 x 20 (a breadth-first
    level of a grid holds at most one anti-diagonal); isolated vertices and
    several components are all placed."""
    rng = np.random.default_rng(9)
    n = 200
    path = sp.diags([np.ones(n - 1), np.ones(n - 1)], [-1, 1], format="csr")
    g = sp.kron(sp.identity(20), sp.diags([np.ones(29), np.ones(29)], [-1, 1])) + \
        sp.kron(sp.diags([np.ones(19), np.ones(19)], [-1, 1]), sp.identity(30))
    two = sp.block_diag([path, sp.csr_matrix((7, 7)), g]).tocsr()
    for S, bound in ((path, 1), (sp.csr_matrix(g), 40), (two, 40)):
        p = rng.permutation(S.shape[0])
        Sp = sp.csr_matrix(S)[p][:, p]
        o = G.rcm_order(_host(Sp))
        assert sorted(o.tolist()) == list(range(1, S.shape[0] + 1))
        assert _bandwidth(Sp, o) <= bound
        assert _bandwidth(Sp, o) < _bandwidth(Sp, np.arange(1, S.shape[0] + 1))


def test_sell_relabel_rcm_ties():
    """GG_SELL_RCM keeps the degree relabelling (slot lengths fall, vinv is a
    permutation, every row kept) and lays rows of equal length out in the
    order of gg_rcm_order; it needs relabelling."""
    R = G.rmat(12, seed=6, dtype=np.float32, values=G.UNIFORM01)
    A = R.to_sell(CPU, 64, 1, 40, heavy_seg_len=64, relabel=True, rcm=True)
    o = G.rcm_order(R)
    rank = np.empty(R.m, np.int64)
    rank[o - 1] = np.arange(R.m)
    rows = [p for p in A.perm if p]
    lens = [l for l, p in zip(A.slot_len, A.perm) if p]
    assert lens == sorted(lens, reverse=True)
    for a, b, la, lb in zip(rows, rows[1:], lens, lens[1:]):
        if la == lb:
            assert rank[a - 1] < rank[b - 1]
    assert len(set(int(v) for v in A.vinv)) == R.m   # distinct internal ids
    S = R.to_scipy()
    for i, j in [(1, 1), (5, int(S.indices[S.indptr[4]]) + 1), (R.m, R.m)]:
        assert A[i, j] == S[i - 1, j - 1]
    with pytest.raises(gg.ArgumentError):
        R.to_sell(CPU, 64, 1, 40, heavy_seg_len=64, relabel=False, rcm=True)


def test_sell_relabel_needs_square():
    S = sp.random(30, 40, density=0.2, format="csr", random_state=np.random.default_rng(1),
                  dtype=np.float32)
    with pytest.raises(gg.ArgumentError):
        gg.SparseGPUMatrixSELL(S, CPU, 64, relabel=True)


def _walk_sell_indexing(A):
    """Every array index the SELL kernels form (sell_slot / heavy segments),
    checked against the buffer lengths on the host."""
    import ctypes as C
    d = A.descriptor(host_ok=True)
    assert gg.lib().gg_matrix_check(C.byref(d)) == 0, gg.lib().gg_last_error()
    n_slots = C.c_int64()
    assert gg.lib().gg_matrix_slots(C.byref(d), C.byref(n_slots)) == 0
    n_slots = n_slots.value
    Cs = A.slice_size
    sp_ = np.asarray(A.slice_ptr)
    perm = None if not d.perm else np.asarray(A.perm)
    slen = None if not d.slot_len else np.asarray(A.slot_len)
    assert n_slots <= A.nslices * Cs
    if perm is not None:
        assert len(perm) == n_slots and len(slen) == n_slots
    for slot in range(n_slots):
        sl, r = divmod(slot, Cs)
        assert sl + 1 < len(sp_)                       # slice_ptr[sl], slice_ptr[sl+1]
        row = (perm[slot] - 1) if perm is not None else slot
        if perm is not None and perm[slot] == 0:
            continue
        assert 0 <= row < A.m
        width = (sp_[sl + 1] - sp_[sl]) // Cs
        ln = slen[slot] if slen is not None else width
        assert ln <= width
        if ln:
            assert sp_[sl] - 1 + r + (ln - 1) * Cs < A.n_stored
    if A.n_heavy:
        hseg = np.asarray(A.heavy_seg)
        hptr = np.asarray(A.heavy_ptr)
        assert len(hseg) == A.n_heavy + 1 and hseg[-1] == A.n_seg == d.heavy_nseg
        assert hptr[-1] == len(A.heavy_col)
        for g in range(A.n_seg):
            h = np.searchsorted(hseg, g, side="right") - 1
            assert 0 <= h < A.n_heavy
            s = hptr[h] + (g - hseg[h]) * A.heavy_seg_len
            assert s < hptr[h + 1]


@pytest.mark.parametrize("case", ["all_heavy", "mixed", "reference", "reference_c64", "empty_rows"])
def test_sell_descriptor_slots_and_indexing(case):
    """Regression: a hybrid whose rows are all heavy has no slots (its empty
    permutation must not read as the identity layout)."""
    rng = np.random.default_rng(3)
    if case == "all_heavy":
        S = sp.random(1000, 1000, density=0.2, format="csr", random_state=rng, dtype=np.float32)
        A = gg.SparseGPUMatrixSELL(S, CPU, 64, sigma=64, heavy_threshold=8, heavy_seg_len=16)
        assert A.nslices == 0 and A.n_heavy == 1000
    elif case == "mixed":
        S = G.rmat(11, seed=2, dtype=np.float32, values=G.UNIFORM01).to_scipy()
        A = gg.SparseGPUMatrixSELL(S, CPU, 64, sigma=64, heavy_threshold=8, heavy_seg_len=16)
        assert A.nslices > 0 and A.n_heavy > 0
    elif case == "reference":
        S = sp.random(300, 300, density=0.05, format="csr", random_state=rng, dtype=np.float32)
        A = gg.SparseGPUMatrixSELL(S, CPU, 32)
    elif case == "reference_c64":
        S = sp.random(1000, 1000, density=0.01, format="csr", random_state=rng, dtype=np.float32)
        A = gg.SparseGPUMatrixSELL(S, CPU, 64)
    else:
        S = sp.csr_matrix((500, 500), dtype=np.float32)
        S[3, 4] = 1.0
        A = gg.SparseGPUMatrixSELL(S, CPU, 64, sigma=64, heavy_threshold=8)
    _walk_sell_indexing(A)


def test_descriptor_check_rejects_uncovered_identity_layout():
    import ctypes as C
    from gpugraphs import _lib as L
    d = L.GGMatrix()
    d.format, d.val_dtype, d.m, d.n = L.SELL, L.FLOAT32, 100, 100
    d.slice_size, d.nslices = 32, 2      # 64 slots for 100 rows
    d.ptr = d.idx = d.val = 8
    assert gg.lib().gg_matrix_check(C.byref(d)) == 2
    d.nslices = 4
    assert gg.lib().gg_matrix_check(C.byref(d)) == 0


def test_rmat_generator_properties():
    R = G.rmat(14, edge_factor=16, seed=7)
    S = R.to_scipy()
    assert R.m == 2**14
    assert (S != S.T).nnz == 0                       # symmetrised
    assert S.diagonal().sum() == 0                   # no self-loops
    assert all(np.all(np.diff(R.colval[R.rowptr[v] - 1:R.rowptr[v + 1] - 1]) > 0)
               for v in range(0, R.m, 97))           # sorted, no duplicates
    deg = R.degrees()
    assert deg.max() > 20 * deg.mean()               # power law (a=.57)
    R2 = G.rmat(14, edge_factor=16, seed=7)
    assert np.array_equal(R.colval, R2.colval)       # deterministic
    assert not np.array_equal(G.rmat(14, seed=8).colval[:1000], R.colval[:1000])


def test_grid_generator_properties():
    g = G.grid(50, 40, long_range_frac=0.01, wmin=1, wmax=1000, seed=3)
    S = g.to_scipy()
    und = 50 * 39 + 40 * 49
    lr = round(0.01 * und)
    assert S.nnz <= 2 * (und + lr) and S.nnz >= 2 * und
    assert S.data.min() >= 1 and S.data.max() <= 1000
    assert np.all(S.data == np.round(S.data))
    P = (S != 0).astype(np.int8)
    assert (P != P.T).nnz == 0                        # both directions present
    assert not np.array_equal(S.data, S.T.tocsr().data)  # independent weights per direction


def test_uniform_rows_generator():
    u = G.uniform_rows(4096, 16, seed=9)
    assert np.all(np.diff(u.rowptr) == 16)
    for v in range(0, 4096, 101):
        c = u.colval[v * 16:(v + 1) * 16]
        assert np.all(np.diff(c) > 0) and c.min() >= 1 and c.max() <= 4096
    assert u.nzval.min() >= 0 and u.nzval.max() < 1


def test_algebra_identities():
    assert gg.monoid_neutral(np.float32, gg.GPUGraphs_min) == np.inf
    assert gg.monoid_absorb(np.int32, gg.GPUGraphs_min) == np.iinfo(np.int32).min
    assert gg.monoid_neutral(np.int32, gg.GPUGraphs_band) == -1
    assert gg.monoid_absorb(np.bool_, gg.GPUGraphs_bor) is True
    assert gg.monoid_absorb(np.int32, gg.GPUGraphs_any) is None
    imax = np.iinfo(np.int32).max
    assert gg.GPUGraphs_safe_add(np.int32(5), np.int32(imax)) == imax
    assert gg.GPUGraphs_safe_add(np.float32(2), np.float32(3)) == 5
    assert gg.GPUGraphs_secondi(0, 1, 1, 1, 7, 1) == 7
    with pytest.raises(gg.MethodError):
        from gpugraphs.algebra import op_code
        op_code(lambda x, y: x)


def test_kernels_refuse_host_arrays():
    A = gg.SparseGPUMatrixCSR(np.eye(4, dtype=np.float32), CPU)
    with pytest.raises(RuntimeError):
        gg.gpu_spmv_(np.zeros(4, np.float32), A, np.ones(4, np.float32))


def test_pack_graphs_block_diagonal():
    """gg_pack_graphs: block g holds graph g's rows with columns shifted by its
    offset; blocks start at multiples of 256; padding rows are empty."""
    gs = [G.rmat(8, seed=1), G.random_graph(300, 4, seed=2),
          G.grid(5, 7, 0.0, 1, 1, seed=3, dtype=np.bool_), G.rmat(9, seed=4)]
    P, off = G.pack(gs)
    assert len(off) == len(gs) + 1 and off[0] == 0 and off[-1] == P.m == P.n
    assert all(o % 256 == 0 for o in off)
    assert P.nnz == sum(g.nnz for g in gs)
    S = P.to_scipy()
    for g, H in enumerate(gs):
        a, b = off[g], off[g + 1]
        blk = S[a:a + H.m, a:a + H.m]
        assert (blk != H.to_scipy()).nnz == 0
        assert S[a:b].nnz == H.nnz                      # nothing outside the block
        assert S[a + H.m:b].nnz == 0                    # padding rows are empty
    with pytest.raises(gg.ArgumentError):
        G.pack([G.rmat(6, seed=1), G.rmat(6, seed=1, dtype=np.int32)])


def test_autotune_candidates_and_signature():
    from gpugraphs import autotune as AT
    R = G.rmat(10, seed=2)
    cands = AT.default_candidates(R)
    assert any(c["layout"] == "csr" for c in cands) and any(c.get("relabel") for c in cands)
    mx = int(R.degrees().max())
    assert all(c.get("heavy", 0) < mx for c in cands)
    assert AT.signature(R, "bfs") != AT.signature(R, "spmv")
    assert AT.signature(R, "bfs") == AT.signature(G.rmat(10, seed=2), "bfs")
    rect = G.HostCSR(3, 5, np.array([1, 2, 3, 4], np.int32), np.array([1, 2, 5], np.int32),
                     np.ones(3, np.float32))
    assert not any(c.get("relabel") for c in AT.default_candidates(rect))
    with pytest.raises(gg.ArgumentError):
        gg.lib()  # library present
        AT.Autotuner(CPU).tune_layout(R, "nope")


def _csr_of(H):
    d = H.nzval.view(np.bool_) if H.nzval.dtype == np.uint8 else H.nzval
    return sp.csr_matrix((d, H.colval - 1, H.rowptr - 1), shape=(H.m, H.n))


@pytest.mark.parametrize("field,symmetry", [("real", "general"), ("integer", "general"),
                                            ("pattern", "general"), ("real", "symmetric"),
                                            ("integer", "skew-symmetric"),
                                            ("pattern", "symmetric")])
def test_read_matrix_market_matches_scipy(tmp_path, field, symmetry):
    """The MatrixMarket reader against scipy.io.mmread on the same file
    (scipy writes it): same pattern, values and type, both triangles for
    symmetric files, transpose on request."""
    import scipy.io
    rng = np.random.default_rng(hash((field, symmetry)) % 2**32)
    n = 40
    S = sp.random(n, n, density=0.15, format="coo", random_state=rng)
    if field == "integer":
        S.data = rng.integers(-50, 50, S.nnz).astype(np.float64)
        S.data[S.data == 0] = 7
    if symmetry == "symmetric":
        S = sp.triu(S) + sp.triu(S, 1).T
    elif symmetry == "skew-symmetric":
        S = sp.triu(S, 1) - sp.triu(S, 1).T
    path = str(tmp_path / "m.mtx")
    scipy.io.mmwrite(path, S.tocoo(), field=field, symmetry=symmetry)
    ref = scipy.io.mmread(path).tocsr()
    ref.sum_duplicates()
    ref.sort_indices()
    H = G.read_matrix_market(path)
    assert H.symmetric == (symmetry != "general")
    want = {"real": np.float64, "integer": np.int64, "pattern": np.bool_}[field]
    got = _csr_of(H)
    assert got.dtype == want
    assert (got != ref.astype(want)).nnz == 0
    np.testing.assert_array_equal(H.rowptr - 1, ref.indptr)
    np.testing.assert_array_equal(H.colval - 1, ref.indices)
    Ht = G.read_matrix_market(path, dtype=np.float32, transpose=True)
    assert Ht.nzval.dtype == np.float32
    assert abs(_csr_of(Ht) - ref.T.astype(np.float32)).max() == 0


def test_read_matrix_market_duplicates_and_errors(tmp_path):
    p = tmp_path / "d.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n% comment\n3 4 4\n"
                 "1 2 1.5\n3 4 2\n1 2 0.25\n2 1 -1\n")
    H = G.read_matrix_market(str(p), dtype=np.float32)
    np.testing.assert_array_equal(_csr_of(H).toarray(),
                                  [[0, 1.75, 0, 0], [-1, 0, 0, 0], [0, 0, 0, 2]])
    bad = tmp_path / "b.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n")
    with pytest.raises(gg.BoundsError):
        G.read_matrix_market(str(bad))
    short = tmp_path / "s.mtx"
    short.write_text("%%MatrixMarket matrix coordinate pattern general\n2 2 3\n1 1\n2 2\n")
    with pytest.raises(gg.ArgumentError):
        G.read_matrix_market(str(short))
    frac = tmp_path / "f.mtx"
    frac.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 0.5\n")
    with pytest.raises(gg.ArgumentError):
        G.read_matrix_market(str(frac), dtype=np.int32)
    with pytest.raises(gg.ArgumentError):
        G.read_matrix_market(str(tmp_path / "missing.mtx"))


@pytest.mark.parametrize("directed", [False, True])
def test_read_edge_list_builds_transpose(tmp_path, directed):
    """Edge list -> A_T (row v lists the tails of the edges into v): the same
    matrix as a COO build of the edges, smallest weight kept on repeats."""
    rng = np.random.default_rng(3 + directed)
    n, e = 50, 300
    u = rng.integers(0, n, e)
    v = rng.integers(0, n, e)
    w = rng.integers(1, 1000, e)
    p = tmp_path / "g.txt"
    p.write_text("# u v w\n" + "".join(f"{a} {b} {c}\n" for a, b, c in zip(u, v, w)))
    H = G.read_edge_list(str(p), base=0, directed=directed, dtype=np.int32)
    assert H.symmetric == (not directed)
    best = {}
    for a, b, c in zip(u, v, w):
        pairs = [(b, a)] if directed or a == b else [(b, a), (a, b)]
        for r, col in pairs:
            best[(r, col)] = min(best.get((r, col), 1 << 30), c)
    m = max(int(u.max()), int(v.max())) + 1
    ref = np.zeros((m, m), np.int64)
    for (r, col), c in best.items():
        ref[r, col] = c
    assert H.m == H.n == m
    np.testing.assert_array_equal(_csr_of(H).toarray(), ref)
    q = tmp_path / "g1.txt"
    q.write_text("1 2\n2 3\n% comment\n3 1\n")
    B = G.read_edge_list(str(q))
    assert B.nzval.dtype == np.bool_ and B.m == 3 and B.nnz == 6
    r = tmp_path / "g2.txt"
    r.write_text("1 2 5\n2 3\n")
    with pytest.raises(gg.ArgumentError):
        G.read_edge_list(str(r), dtype=np.int32)


def test_tune_default_candidates_native():
    """The native candidate set: CSR lane counts, relabelled SELL only for a
    square matrix, heavy thresholds only below the longest row, and the
    sigma-4096 window layout."""
    from gpugraphs import autotune as AT, generators as GEN
    R = GEN.rmat(10, seed=3)
    cands = AT.default_candidates(R)
    mx = int(R.degrees().max())
    assert [c["lanes"] for c in cands if c["layout"] == "csr"] == [1, 4, 16, 64]
    sells = [c for c in cands if c["layout"] == "sell"]
    assert all(c["heavy"] < mx for c in sells)
    assert any(c["sigma"] == 4096 and not c["relabel"] for c in sells)
    assert any(c["relabel"] for c in sells)
    rect = GEN.HostCSR(R.m, R.n + 7, R.rowptr, R.colval, R.nzval)
    assert not any(c.get("relabel") for c in AT.default_candidates(rect))
    for c in cands:   # the dict <-> C struct mapping round-trips
        assert AT._from_native(AT._to_native(c)) == c
