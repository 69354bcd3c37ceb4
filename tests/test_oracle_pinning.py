"""Pins the CPU oracle (oracle/gg_oracle.cpp) against checks that do not
restate it: the reference's own known-answer data, library routines (scipy
FP64 products, csgraph BFS / Dijkstra), brute force on tiny graphs and
closed-form special cases of src/algebra.jl.  CPU only."""
import itertools
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse import csgraph

from conftest import GOLDEN, random_csr
import graphs as G

INT32_MAX = np.iinfo(np.int32).max


def test_sell_known_answer(oracle):
    """test/structs.jl:176-190: slice_ptr / colval / nzval of the 4x4 KAT."""
    kat = json.load(open(os.path.join(GOLDEN, "sell_kat.json")))
    A = sp.csr_matrix(np.array(kat["dense"], np.float32))
    S = oracle.sell_from_csr(oracle.from_scipy_csr(A), kat["slice_size_arg"])
    e = kat["expected"]
    assert S.slice_size == e["slice_size"] and S.nslices == e["nslices"]
    assert S.n_stored == e["n_stored"] and S.nnz == e["nnz"]
    assert S.ptr.tolist() == e["slice_ptr"]
    assert S.idx.tolist() == e["colval"]
    assert S.val.tolist() == e["nzval"]


def test_csr_constructor_errors(oracle):
    """test/structs.jl:79-150: every malformed triple is an ArgumentError."""
    cases = json.load(open(os.path.join(GOLDEN, "csr_errors.json")))["cases"]
    for c in cases:
        s = oracle.csr_check(c["m"], c["n"], c["rowptr"], c["colval"], c["nzval_len"])
        assert oracle.STATUS[s] == c["expect"], c["what"]


@pytest.mark.parametrize("m,n,p", [(10, 10, 0.5), (1000, 1000, 0.2), (300, 700, 0.05),
                                   (700, 300, 0.05)])
@pytest.mark.parametrize("fmt", ["csr", "sell"])
def test_plus_times_spmv_vs_fp64(oracle, m, n, p, fmt):
    """Plus-times SpMV vs scipy's FP64 product (test/spmv.jl:10-70).  The
    non-square shapes catch a transposed operand."""
    rng = np.random.default_rng(m * 7 + n)
    S = random_csr(rng, m, n, p)
    x = rng.random(n, dtype=np.float32)
    ref = S.astype(np.float64) @ x.astype(np.float64)
    A = oracle.from_scipy_csr(S)
    if fmt == "sell":
        if m > n:
            pytest.skip("reference SELL kernel guards rows with A.n (src/spmv/SELL.jl:50)")
        A = oracle.sell_from_csr(A, 32)
    c = np.zeros(m, np.float32)
    oracle.spmv(c, A, x)
    np.testing.assert_allclose(c, ref, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("fmt", ["csr", "sell"])
def test_masked_spmv_closed_form(oracle, fmt):
    """test/spmv.jl:72-170: C = A*B .* mask with unmasked rows left at their prior value."""
    rng = np.random.default_rng(5)
    S = random_csr(rng, 1000, 1000, 0.2)
    x = rng.random(1000, dtype=np.float32)
    mask = rng.random(1000) < 0.5
    prior = np.full(1000, -7.0, np.float32)
    ref = np.where(mask, S.astype(np.float64) @ x, -7.0)
    A = oracle.from_scipy_csr(S)
    if fmt == "sell":
        A = oracle.sell_from_csr(A, 32)
    c = prior.copy()
    oracle.spmv(c, A, x, mask=mask.astype(np.int32))
    np.testing.assert_allclose(c, ref, rtol=1e-5, atol=1e-6)
    # a range restricts rows the same way (src/spmv/CSR.jl:224-248)
    c2 = prior.copy()
    oracle.spmv(c2, A, x, range_=(101, 200))
    ref2 = prior.astype(np.float64).copy()
    ref2[100:300] = (S.astype(np.float64) @ x)[100:300]
    np.testing.assert_allclose(c2, ref2, rtol=1e-5, atol=1e-6)


def test_csc_spmv_vs_fp64(oracle):
    rng = np.random.default_rng(9)
    S = random_csr(rng, 400, 400, 0.05).tocsc()
    S.sort_indices()
    x = rng.random(400, dtype=np.float32)
    A = oracle.csc(400, 400, S.indptr + 1, S.indices + 1, S.data)
    c = np.zeros(400, np.float32)
    oracle.spmv(c, A, x)
    np.testing.assert_allclose(c, S.astype(np.float64) @ x, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("k", [5, 20])
@pytest.mark.parametrize("fmt", ["csr", "sell"])
def test_spmm_vs_fp64(oracle, k, fmt):
    """test/spmm.jl:10-73."""
    rng = np.random.default_rng(k)
    S = random_csr(rng, 300, 300, 0.1)
    B = np.asfortranarray(rng.random((300, k), dtype=np.float32))
    A = oracle.from_scipy_csr(S)
    if fmt == "sell":
        A = oracle.sell_from_csr(A, 32)
    Cm = np.zeros((300, k), np.float32, order="F")
    oracle.spmm(Cm, A, B)
    np.testing.assert_allclose(Cm, S.astype(np.float64) @ B.astype(np.float64), rtol=1e-5,
                               atol=1e-6)


def test_algebra_special_cases(oracle):
    """Closed forms of src/algebra.jl: safe_add saturates at typemax
    (:47-48), the CSR kernel stops at the absorbing element (src/spmv/CSR.jl:49-51)
    while the SELL SpMV kernel has no such exit (src/spmv/SELL.jl:54-60), and
    GPUGraphs_any with a mask takes the first frontier entry on CSR
    (src/spmv/CSR.jl:159-170) but the largest on SELL (max, algebra.jl:35)."""
    # safe_add with typemax: min over {typemax+w} stays typemax, finite entries add
    A = oracle.csr(2, 3, [1, 3, 4], [1, 2, 3], np.array([5, 7, 1], np.int32))
    x = np.array([INT32_MAX, 3, INT32_MAX], np.int32)
    c = np.full(2, INT32_MAX, np.int32)
    oracle.spmv(c, A, x, add=oracle.OP_MIN, mul=oracle.OP_SAFE_ADD, accum=oracle.OP_MIN)
    assert c.tolist() == [10, INT32_MAX]
    # integer plus: CSR stops once acc == typemax; SELL keeps wrapping
    A = oracle.csr(1, 2, [1, 3], [1, 2], np.array([INT32_MAX, 5], np.int32))
    x = np.array([1, 1], np.int32)
    c = np.zeros(1, np.int32)
    oracle.spmv(c, A, x)
    assert c[0] == INT32_MAX
    c = np.zeros(1, np.int32)
    oracle.spmv(c, oracle.sell_from_csr(A, 32), x)
    assert c[0] == np.int32(INT32_MAX + 5 - 2**32)
    # any.secondi: CSR+mask -> first frontier column, SELL -> largest
    A = oracle.csr(1, 4, [1, 4], [2, 3, 4], np.array([1, 1, 1], np.int32))
    curr = np.array([0, 1, 0, 1], np.int32)
    mask = np.array([1], np.int32)
    c = np.zeros(1, np.int32)
    oracle.spmv(c, A, curr, add=oracle.OP_ANY, mul=oracle.OP_SECONDI, mask=mask)
    assert c[0] == 2
    c = np.zeros(1, np.int32)
    oracle.spmv(c, oracle.sell_from_csr(A, 32), curr, add=oracle.OP_ANY, mul=oracle.OP_SECONDI,
                mask=mask)
    assert c[0] == 4
    # range + mask on CSR takes the generic kernel: any = max (src/spmv/CSR.jl:338-353)
    c = np.zeros(1, np.int32)
    oracle.spmv(c, A, curr, add=oracle.OP_ANY, mul=oracle.OP_SECONDI, mask=mask, range_=(1, 1))
    assert c[0] == 4
    # or-and on Bool, absorbing true
    A = oracle.csr(2, 2, [1, 3, 3], [1, 2], np.array([1, 1], np.uint8))
    c = np.zeros(2, np.uint8)
    oracle.spmv(c, A, np.array([0, 1], np.uint8), add=oracle.OP_BOR, mul=oracle.OP_BAND)
    assert c.tolist() == [1, 0]


def test_argument_errors(oracle):
    """_validate_args order, src/spmv/CSR.jl:374-402."""
    A = oracle.csr(3, 4, [1, 2, 3, 4], [1, 2, 3], np.ones(3, np.float32))
    with pytest.raises(oracle.OracleError, match="DimensionMismatch"):
        oracle.spmv(np.zeros(3, np.float32), A, np.ones(5, np.float32))
    with pytest.raises(oracle.OracleError, match="DimensionMismatch"):
        oracle.spmv(np.zeros(4, np.float32), A, np.ones(4, np.float32))
    with pytest.raises(oracle.OracleError, match="ArgumentError"):
        oracle.spmv(np.zeros(3, np.int32), A, np.ones(4, np.float32))
    with pytest.raises(oracle.OracleError, match="DimensionMismatch"):
        oracle.spmv(np.zeros(3, np.float32), A, np.ones(4, np.float32),
                    mask=np.ones(2, np.int32))


def _graphs():
    out = []
    for seed in (1, 2):
        out.append(("ba", 30, G.barabasi_albert(30, 3, seed)))
        out.append(("dm", 30, G.dorogovtsev_mendes(30, seed)))
    out.append(("ba300", 300, G.barabasi_albert(300, 2, 3)))
    return out


@pytest.mark.parametrize("name,n,edges", _graphs())
@pytest.mark.parametrize("dtype", [np.bool_, np.int32])
@pytest.mark.parametrize("fmt", ["csr", "sell"])
def test_bfs_distances_vs_csgraph(oracle, name, n, edges, dtype, fmt):
    """test/bfs.jl:10-46: levels equal Graphs.gdistances; here csgraph BFS."""
    Adj = G.undirected_adj(n, edges, dtype)
    A = oracle.from_scipy_csr(Adj.T.tocsr())
    if fmt == "sell":
        A = oracle.sell_from_csr(A, 32)
    for s in (1, n // 2, n):
        ref = csgraph.shortest_path(Adj.astype(np.float64), directed=True, unweighted=True,
                                    indices=s - 1)
        ref = np.where(np.isinf(ref), INT32_MAX, ref).astype(np.int32)
        for use_mask in ((True, False) if dtype == np.int32 else (True,)):
            d = oracle.bfs_distances(A, s, use_mask=use_mask)
            np.testing.assert_array_equal(d, ref)


def test_bfs_directed_with_unreachable(oracle):
    rng = np.random.default_rng(11)
    Adj = G.random_digraph(rng, 200, 0.012, dtype=np.int32)
    A = oracle.from_scipy_csr(Adj.T.tocsr())
    for s in (1, 50, 199):
        ref = csgraph.shortest_path(Adj.astype(np.float64), directed=True, unweighted=True,
                                    indices=s - 1)
        ref = np.where(np.isinf(ref), INT32_MAX, ref).astype(np.int32)
        assert (ref == INT32_MAX).any()
        np.testing.assert_array_equal(oracle.bfs_distances(A, s), ref)


def _brute_levels(n, edges_dir, s):
    """Hop distances by exhaustive relaxation over all vertex orders (n <= 7)."""
    INF = INT32_MAX
    best = [INF] * n
    best[s] = 0
    adj = {u: [v for (a, v) in edges_dir if a == u] for u in range(n)}
    # enumerate every simple path from s by DFS
    def dfs(u, d, seen):
        for v in adj[u]:
            if v not in seen:
                best[v] = min(best[v], d + 1)
                dfs(v, d + 1, seen | {v})
    dfs(s, 0, {s})
    return best


def test_bfs_brute_force_tiny(oracle):
    rng = np.random.default_rng(3)
    for trial in range(40):
        n = int(rng.integers(1, 7))
        pairs = [(u, v) for u in range(n) for v in range(n) if u != v and rng.random() < 0.35]
        M = np.zeros((n, n), np.int32)
        for u, v in pairs:
            M[u, v] = 1
        A = oracle.from_scipy_csr(sp.csr_matrix(M.T))
        s = int(rng.integers(n)) + 1
        assert oracle.bfs_distances(A, s).tolist() == _brute_levels(n, pairs, s - 1)


def _check_parents_tree(Adj, s, par, dist, pick):
    """parents form a BFS tree: parent is an in-neighbour one level up, source
    0, unreached typemax; pick names the reference's tie-break."""
    AT = Adj.T.tocsr()
    for v in range(Adj.shape[0]):
        if v == s - 1:
            assert par[v] == 0
        elif dist[v] == INT32_MAX:
            assert par[v] == INT32_MAX
        else:
            nb = AT.indices[AT.indptr[v]:AT.indptr[v + 1]]
            cand = [u + 1 for u in nb if dist[u] == dist[v] - 1]
            assert par[v] in cand
            if pick == "min":
                assert par[v] == min(cand)
            elif pick == "max":
                assert par[v] == max(cand)


@pytest.mark.parametrize("fmt,pick", [("csr", "min"), ("sell", "max")])
def test_bfs_parents_valid_tree(oracle, fmt, pick):
    """bfs_parents (src/algorithms/bfs.jl:113-183) on Int32 adjacency; Bool
    adjacency throws in the reference (src/spmv/CSR.jl:381-387)."""
    for name, n, edges in _graphs():
        Adj = G.undirected_adj(n, edges, np.int32)
        A = oracle.from_scipy_csr(Adj.T.tocsr())
        if fmt == "sell":
            A = oracle.sell_from_csr(A, 32)
        for s in (1, n):
            dist = oracle.bfs_distances(A, s)
            for batch in (1, 3, 8):
                par = oracle.bfs_parents(A, s, batch_size=batch)
                _check_parents_tree(Adj, s, par, dist, pick)


@pytest.mark.parametrize("fmt", ["csr", "sell"])
def test_shortest_path_vs_dijkstra(oracle, fmt):
    """test/shortest_path.jl:31-80: integer weights vs Dijkstra (exact)."""
    rng = np.random.default_rng(21)
    for n, p in ((30, 0.1), (250, 0.02)):
        for dtype in (np.float32, np.int32):
            Adj = G.random_digraph(rng, n, p, wmax=1000, dtype=dtype)
            A = oracle.from_scipy_csr(Adj.T.tocsr())
            if fmt == "sell":
                A = oracle.sell_from_csr(A, 32)
            for s in (1, n // 3, n):
                ref = csgraph.dijkstra(Adj.astype(np.float64), directed=True, indices=s - 1)
                d = oracle.shortest_path(A, s)
                if dtype == np.float32:
                    ref = ref.astype(np.float32)
                else:
                    ref = np.where(np.isinf(ref), INT32_MAX, ref).astype(np.int32)
                np.testing.assert_array_equal(d, ref)


def test_shortest_path_bool_is_hops(oracle):
    """Bool weights give Int32 hop counts (src/algorithms/shortest_path.jl:18)."""
    for name, n, edges in _graphs()[:2]:
        Adj = G.undirected_adj(n, edges, np.bool_)
        A = oracle.sell_from_csr(oracle.from_scipy_csr(Adj.T.tocsr()), 32)
        d = oracle.shortest_path(A, 1)
        assert d.dtype == np.int32
        np.testing.assert_array_equal(d, oracle.bfs_distances(A, 1))


@pytest.mark.parametrize("dt", [np.int32, np.int64, np.float32, np.float64])
@pytest.mark.parametrize("fmt", ["csr", "sell"])
def test_shortest_path_multi_all_sources(oracle, fmt, dt):
    """test/shortest_path.jl:82-126: all n sources vs per-source Dijkstra, for
    each distance type (unreached: typemax for integers, Inf for floats)."""
    rng = np.random.default_rng(4)
    Adj = G.random_digraph(rng, 40, 0.08, wmax=10, dtype=np.int32).astype(dt)
    A = oracle.from_scipy_csr(Adj.T.tocsr())
    if fmt == "sell":
        A = oracle.sell_from_csr(A, 32)
    D = oracle.shortest_path_multi(A, np.arange(1, 41))
    assert D.dtype == dt
    ref = csgraph.dijkstra(Adj.astype(np.float64), directed=True).T  # column i = source i
    if np.issubdtype(dt, np.integer):
        fin = ~np.isinf(ref)
        want = np.full(ref.shape, np.iinfo(dt).max, dt)
        want[fin] = ref[fin].astype(dt)
        ref = want
    np.testing.assert_array_equal(D, ref.astype(dt))


def test_spmspv_vs_fp64_and_mask(oracle):
    rng = np.random.default_rng(8)
    S = random_csr(rng, 500, 400, 0.02).tocsc()
    S.sort_indices()
    A = oracle.csc(500, 400, S.indptr + 1, S.indices + 1, S.data)
    nz = np.sort(rng.choice(400, 40, replace=False)) + 1
    bv = rng.random(40, dtype=np.float32)
    b = np.zeros(400)
    b[nz - 1] = bv
    mask = (rng.random(500) < 0.5).astype(np.int32)
    for comp in (False, True):
        ci, cv = oracle.spmspv(A, nz, bv, 400, oracle.OP_ADD, oracle.OP_MUL, np.float32,
                               mask=mask, mask_complement=comp)
        keep = (mask == 0) if comp else (mask != 0)
        touched = (abs(S.astype(np.float64)) @ (b != 0).astype(np.float64)) > 0
        exp_idx = np.nonzero(touched & keep)[0] + 1
        np.testing.assert_array_equal(ci, exp_idx)
        ref = S.astype(np.float64) @ b
        np.testing.assert_allclose(cv, ref[ci - 1], rtol=1e-5, atol=1e-7)


def _tiny_csc(rng, m, n, p, dtype):
    S = sp.random(m, n, density=p, format="csc", random_state=rng, dtype=np.float64)
    S.data = rng.integers(1, 9, S.nnz).astype(np.float64)
    S.sort_indices()
    return S.astype(dtype)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_spmspv_semirings_brute_force(oracle, seed):
    """Min-plus, any-secondi and or-and SpMSpV against a dense double loop over
    rows and the frontier's columns (no sparse restatement): c_i folds
    mul(A[i,j], b_j, j) over the stored A[i,j] of b's columns, ascending j;
    rows outside the mask (or inside it, for the complement) take nothing."""
    rng = np.random.default_rng(seed)
    m, n = 23, 17
    for add, mul, vt, ot in [(oracle.OP_MIN, oracle.OP_SAFE_ADD, np.float32, np.float32),
                             (oracle.OP_MIN, oracle.OP_SAFE_ADD, np.int32, np.int32),
                             (oracle.OP_ANY, oracle.OP_SECONDI, np.int32, np.int32),
                             (oracle.OP_BOR, oracle.OP_BAND, np.bool_, np.bool_)]:
        S = _tiny_csc(rng, m, n, 0.3, vt)
        D = S.toarray()
        stored = (S != 0).toarray() if vt != np.bool_ else S.toarray()
        A = oracle.csc(m, n, S.indptr + 1, S.indices + 1, S.data)
        nz = np.sort(rng.choice(n, 6, replace=False)) + 1
        if vt == np.bool_:
            bv = np.ones(6, np.bool_)
        elif mul == oracle.OP_SECONDI:
            bv = np.ones(6, np.int32)
        else:
            bv = rng.integers(0, 50, 6).astype(vt)
        mask = (rng.random(m) < 0.6).astype(np.int32)
        for use_mask, comp in ((False, False), (True, False), (True, True)):
            ci, cv = oracle.spmspv(A, nz, bv, n, add, mul, ot,
                                   mask=mask if use_mask else None, mask_complement=comp)
            exp_i, exp_v = [], []
            for i in range(m):
                if use_mask and (mask[i] != 0) == comp:
                    continue
                vals = [(j, D[i, j - 1], b) for j, b in zip(nz, bv) if stored[i, j - 1]]
                if not vals:
                    continue
                if add == oracle.OP_MIN:
                    acc = min(a + b for _, a, b in vals)
                elif add == oracle.OP_ANY:
                    acc = max(j * b for j, _, b in vals)   # any = max, secondi = col * b
                else:
                    acc = any(bool(a) and bool(b) for _, a, b in vals)
                exp_i.append(i + 1)
                exp_v.append(acc)
            np.testing.assert_array_equal(ci, np.array(exp_i, np.int32))
            np.testing.assert_array_equal(cv, np.array(exp_v).astype(ot))


@pytest.mark.parametrize("fmt", ["csr", "sell"])
def test_spmm_mask_and_range_closed_form(oracle, fmt):
    """Row mask and row range of gpu_spmm! (src/spmm/CSR.jl:36-109):
    selected rows get A*B, every other row of C keeps its prior value."""
    rng = np.random.default_rng(17)
    S = random_csr(rng, 200, 200, 0.1)
    k = 7
    B = np.asfortranarray(rng.random((200, k), dtype=np.float32))
    full = S.astype(np.float64) @ B.astype(np.float64)
    A = oracle.from_scipy_csr(S)
    if fmt == "sell":
        A = oracle.sell_from_csr(A, 32)
    prior = np.asfortranarray(np.full((200, k), -3.0, np.float32))
    if fmt == "csr":
        mask = (rng.random(200) < 0.4).astype(np.int32)
        Cm = prior.copy(order="F")
        oracle.spmm(Cm, A, B, mask=mask)
        want = np.where(mask[:, None] != 0, full, -3.0)
        np.testing.assert_allclose(Cm, want, rtol=1e-5, atol=1e-6)
        Cr = prior.copy(order="F")
        oracle.spmm(Cr, A, B, range_=(41, 100))
        want = np.full((200, k), -3.0)
        want[40:140] = full[40:140]
        np.testing.assert_allclose(Cr, want, rtol=1e-5, atol=1e-6)
    else:
        # SELL SpMM has neither mask nor range (src/spmm/SELL.jl:84-102)
        with pytest.raises(oracle.OracleError):
            oracle.spmm(prior.copy(order="F"), A, B, mask=np.ones(200, np.int32))


@pytest.mark.parametrize("name,n,edges", _graphs())
def test_bfs_queue_levels_vs_csgraph_and_pull_loop(oracle, name, n, edges):
    """The full-size BFS checker (textbook queue BFS) against scipy's BFS and
    against the restated pull loop, from every source of a small graph."""
    Adj = G.undirected_adj(n, edges, np.bool_)
    A = oracle.from_scipy_csr(Adj.T.tocsr())
    D = oracle.bfs_levels_queue(A, np.arange(1, n + 1))
    ref = csgraph.shortest_path(Adj.astype(np.float64), unweighted=True, directed=True)
    for s in range(n):
        r = ref[s]
        want = np.where(np.isinf(r), INT32_MAX, r).astype(np.int32)
        np.testing.assert_array_equal(D[:, s], want)
        np.testing.assert_array_equal(D[:, s], oracle.bfs_distances(A, s + 1))


def test_bfs_queue_directed_and_even_weights(oracle):
    """Directed digraph with Int32 values: an even value is an inactive edge
    under band with the frontier's ones, as in the pull loop."""
    rng = np.random.default_rng(3)
    Adj = G.random_digraph(rng, 300, 0.01, wmax=6, dtype=np.int32)
    AT = Adj.T.tocsr()
    A = oracle.from_scipy_csr(AT)
    srcs = [1, 50, 300]
    D = oracle.bfs_levels_queue(A, srcs)
    odd = Adj.copy()
    odd.data = (odd.data % 2 == 1).astype(np.float64)
    odd.eliminate_zeros()
    ref = csgraph.shortest_path(odd.astype(np.float64), unweighted=True, directed=True)
    for q, s in enumerate(srcs):
        want = np.where(np.isinf(ref[s - 1]), INT32_MAX, ref[s - 1]).astype(np.int32)
        np.testing.assert_array_equal(D[:, q], want)
        np.testing.assert_array_equal(D[:, q], oracle.bfs_distances(A, s))


@pytest.mark.parametrize("dt", [np.float32, np.int32])
def test_sssp_certificate_accepts_exact_and_rejects_mistakes(oracle, dt):
    """The full-size Bellman-Ford checker: Dijkstra's distances pass; a
    distance one too high or one too low, a source not at zero, a reachable
    vertex marked unreached and an unreachable one given a value all fail."""
    rng = np.random.default_rng(11)
    Adj = G.random_digraph(rng, 400, 0.008, wmax=1000, dtype=np.float32)
    AT = Adj.T.tocsr().astype(dt)
    A = oracle.from_scipy_csr(AT)
    srcs = [1, 200]
    ref = csgraph.dijkstra(Adj.astype(np.float64), directed=True, indices=[s - 1 for s in srcs]).T
    top = np.iinfo(dt).max if np.issubdtype(dt, np.integer) else np.inf
    D = np.asfortranarray(np.where(np.isinf(ref), top, ref).astype(dt))
    assert oracle.sssp_certificate(A, srcs, D).tolist() == [0, 0]
    assert oracle.sssp_certificate(A, srcs, oracle.shortest_path_multi(A, srcs)).tolist() == [0, 0]
    fin = np.nonzero(np.isfinite(ref[:, 0]))[0]
    fin = fin[fin != 0]
    unr = np.nonzero(np.isinf(ref[:, 0]))[0]
    v = int(fin[len(fin) // 2])
    for change in ("plus", "minus", "source", "drop", "invent"):
        E = D.copy(order="F")
        if change == "plus":
            E[v, 0] += 1
        elif change == "minus":
            E[v, 0] -= 1
        elif change == "source":
            E[0, 0] = 1
        elif change == "drop":
            E[v, 0] = top
        else:
            if len(unr) == 0:
                continue
            E[unr[0], 0] = 5
        assert oracle.sssp_certificate(A, srcs, E)[0] > 0, change
