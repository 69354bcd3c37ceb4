"""GPU tests of the continuous-batching server: every query's result equals
the oracle's single-source BFS (src/algorithms/bfs.jl:31-73) / Bellman-Ford
(src/algorithms/shortest_path.jl:34-83) from its source, whether it runs
alone or in a mix that overflows the engines' slots and the memory budget;
a query that can never fit the budget fails with BudgetExceeded and one
whose source reaches a negative cycle with NotConvergedError."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


def _orc(orc, H, pattern=False):
    return orc.csr(H.m, H.n, H.rowptr, H.colval,
                   np.ones(H.nnz, np.uint8) if pattern else H.nzval)


@pytest.fixture(scope="module")
def graphs(oracle):
    import torch
    import gpugraphs as gg
    from gpugraphs import generators as GEN
    assert torch.cuda.is_available()
    R = gg.ROCBackend(0)
    # weights 1..100: SSSP is min-plus, BFS follows the or-and edges (odd values)
    g0 = GEN.random_graph(2048, 8, seed=31, dtype=np.int32, values=GEN.INTEGER, wmin=1, wmax=100)
    g1 = GEN.grid(30, 40, 0.01, 1, 50, seed=5, dtype=np.int32)   # relabelled: sparse SSSP steps
    g2 = GEN.rmat(13, seed=2)
    g3 = GEN.rmat(12, seed=3)   # symmetric pattern, relabelled: sparse SSSP steps
    hosts = [g0, g1, g2, g3]
    mats = [g0.to_csr(R), g1.to_sell(R, 64, 1, 8, 64, relabel=True), g2.to_sell(R, 64, 1, 64, 256),
            g3.to_sell(R, 64, 1, 128, 256, relabel=True)]
    return gg, torch, mats, [_orc(oracle, h) for h in hosts], oracle


def _expected(orc, Ao, kind, src):
    from gpugraphs import serving as S
    return orc.bfs_distances(Ao, src) if kind == S.BFS else orc.shortest_path(Ao, src)


def _queries(rng, mats, count):
    from gpugraphs import serving as S
    q = []
    for _ in range(count):
        g = int(rng.integers(len(mats)))
        q.append((g, S.BFS if rng.random() < 0.5 else S.SSSP, int(rng.integers(mats[g].m)) + 1))
    return q


def test_server_results_equal_alone_and_in_mix(graphs):
    gg, torch, mats, omats, orc = graphs
    from gpugraphs import serving as S
    rng = np.random.default_rng(5)
    qs = _queries(rng, mats, 600)          # > 64 BFS slots / 64 SSSP columns per graph
    exp = {i: _expected(orc, omats[g], k, s) for i, (g, k, s) in enumerate(qs)}
    with S.GraphServer(budget_bytes=64 << 20) as srv:
        gids = [srv.add_graph(A) for A in mats]
        srv.start()
        # alone: one query at a time
        for i in range(6):
            g, k, s = qs[i]
            print_case("serve alone", g, k, s)
            rid = srv.submit(gids[g], k, s)
            np.testing.assert_array_equal(srv.fetch(rid, gids[g], k), exp[i])
        # in a mix: everything at once
        print_case("serve mix", len(qs))
        rids = [srv.submit(gids[g], k, s) for g, k, s in qs]
        for i, rid in enumerate(rids):
            g, k, s = qs[i]
            np.testing.assert_array_equal(srv.fetch(rid, gids[g], k), exp[i],
                                          err_msg=f"query {i}: graph {g} kind {k} source {s}")
        st = srv.stats()
        assert st["completed"] == len(qs) + 6 and st["failed"] == 0
        assert st["arena_used"] == 0
        srv.stop()
        # every finished query is counted by exactly one engine, and a query
        # needs at least one step of its engine
        es = [srv.engine_stats(g, k) for g in gids for k in (S.BFS, S.SSSP)]
        assert sum(e["finished"] for e in es) == len(qs) + 6
        assert sum(e["steps"] for e in es) == st["steps"]
        for e in es:
            assert e["slots"] == 64 and e["busy_slot_steps"] <= 64 * e["steps"]
            assert (e["finished"] == 0) or (e["steps"] > 0 and e["device_s"] > 0.0)
        from gpugraphs._lib import ArgumentError
        with pytest.raises(ArgumentError, match="unknown graph"):
            srv.engine_stats(len(gids), S.BFS)


def test_server_budget_admission_and_rejection(graphs):
    gg, torch, mats, omats, orc = graphs
    from gpugraphs import serving as S
    A, Big = mats[1], mats[2]
    n = A.m
    rng = np.random.default_rng(9)
    srcs = rng.integers(1, n + 1, 40)
    exp = [_expected(orc, omats[1], S.BFS, int(s)) for s in srcs]
    budget = 3 * 4 * n + 256               # room for 3 results; the rest wait for memory
    assert Big.m * 4 > budget
    with S.GraphServer(budget_bytes=budget) as srv:
        gid = srv.add_graph(A)
        big = srv.add_graph(Big)
        srv.start()
        print_case("serve budget", n)
        rids = [srv.submit(gid, S.BFS, int(s)) for s in srcs]
        rid_big = srv.submit(big, S.BFS, 1)
        for rid, e in zip(rids, exp):
            np.testing.assert_array_equal(srv.fetch(rid, gid, S.BFS), e)
        with pytest.raises(gg.BudgetExceeded):
            srv.fetch(rid_big, big, S.BFS)
        assert srv.stats()["failed"] == 1
        srv.stop()


def test_server_top_down_steps_and_slot_reuse(graphs):
    """A symmetric unit pattern with small frontiers (a grid) takes narrow
    top-down steps, which write only the rows they touch; 200 queries through
    64 slots reuse retired slots whose rows may hold stale words.  Every
    result equals the oracle's BFS from its source."""
    gg, torch, mats, omats, orc = graphs
    from gpugraphs import serving as S
    from gpugraphs import generators as GEN
    Hw = GEN.grid(60, 70, 0.0, 1, 1, seed=7, dtype=np.int32)
    Hp = GEN.HostCSR(Hw.m, Hw.n, Hw.rowptr, Hw.colval, np.ones(Hw.nnz, np.bool_), True)
    M = Hp.to_sell(gg.ROCBackend(0), 64, 1, 0, 256, relabel=True)
    Ao = _orc(orc, Hp, pattern=True)
    rng = np.random.default_rng(23)
    srcs = [int(s) for s in rng.integers(1, Hp.m + 1, 200)]
    with S.GraphServer(budget_bytes=256 << 20) as srv:
        gid = srv.add_graph(M)
        srv.start()
        print_case("serve top-down reuse", Hp.m, len(srcs))
        rids = []
        for i, s in enumerate(srcs):   # arrivals spread over the traversals in flight
            rids.append(srv.submit(gid, S.BFS, s))
            if i % 40 == 39:
                srv.fetch(rids[i - 20], gid, S.BFS)   # forces progress between waves
                rids[i - 20] = None
        for i, rid in enumerate(rids):
            if rid is None:
                continue
            np.testing.assert_array_equal(srv.fetch(rid, gid, S.BFS),
                                          orc.bfs_distances(Ao, srcs[i]),
                                          err_msg=f"query {i} from {srcs[i]}")
        srv.stop()


@pytest.mark.parametrize("dt", [np.float64, np.int64, np.float32])
def test_server_sssp_busy_column_groups(graphs, dt):
    """SSSP steps on a SELL engine relax only the groups of 16 bytes of
    columns that hold a busy slot (2 columns of 8-byte distances, 4 of 4-byte
    ones).  180 queries through 64 columns, submitted in waves while earlier
    ones finish, free slots in any order, so the busy groups are scattered
    and change between steps; every result equals the oracle's
    Bellman-Ford from its source."""
    gg, torch, mats, omats, orc = graphs
    from gpugraphs import serving as S
    from gpugraphs import generators as GEN
    H = GEN.grid(37, 29, 0.02, 1, 900, seed=int(np.dtype(dt).itemsize) + 40, dtype=np.int32)
    Hd = GEN.HostCSR(H.m, H.n, H.rowptr, H.colval, H.nzval.astype(dt), True)
    M = Hd.to_sell(gg.ROCBackend(0), 64, 1, 0, 256, relabel=True)
    Ao = _orc(orc, Hd)
    rng = np.random.default_rng(int(np.dtype(dt).itemsize) * 3 + 1)
    srcs = [int(s) for s in rng.integers(1, H.m + 1, 180)]
    with S.GraphServer(budget_bytes=256 << 20) as srv:
        gid = srv.add_graph(M)
        srv.start()
        print_case("serve sssp column groups", np.dtype(dt).name, H.m, len(srcs))
        rids = []
        for i, s in enumerate(srcs):
            rids.append(srv.submit(gid, S.SSSP, s))
            if i % 30 == 29:
                j = i - int(rng.integers(5, 25))
                np.testing.assert_array_equal(srv.fetch(rids[j], gid, S.SSSP),
                                              orc.shortest_path(Ao, srcs[j]),
                                              err_msg=f"query {j} from {srcs[j]}")
                rids[j] = None
        for i, rid in enumerate(rids):
            if rid is None:
                continue
            np.testing.assert_array_equal(srv.fetch(rid, gid, S.SSSP),
                                          orc.shortest_path(Ao, srcs[i]),
                                          err_msg=f"query {i} from {srcs[i]}")
        srv.stop()


@pytest.mark.parametrize("fmt", ["csr", "sell_relabel"])
def test_server_negative_cycle_fails_not_converged(graphs, fmt):
    """A query whose source reaches a negative cycle fails with
    NotConvergedError after m + 16 steps (gg_shortest_path's cap); a query
    from a source that cannot reach the cycle, on the same engine at the
    same time, still returns the oracle's distances.  The failed query's
    column is not extracted (and so not reset by its extraction): 70 later
    queries reuse every slot, the failed one first, and each returns the
    oracle's distances."""
    gg, torch, mats, omats, orc = graphs
    from gpugraphs import serving as S
    import scipy.sparse as sp
    n = 64
    # path 1 -> 2 -> ... -> 40 with weight 3, cycle 10 -> 11 -> 12 -> 10 of
    # total weight -3, and a separate chain 41 -> ... -> 64 with weight 2
    edges = [(i, i + 1, 3) for i in range(40 - 1)] + [(11, 9, -9)]
    edges += [(i, i + 1, 2) for i in range(40, n - 1)]
    u = np.array([e[0] for e in edges])
    v = np.array([e[1] for e in edges])
    w = np.array([e[2] for e in edges], np.int32)
    Adj = sp.csr_matrix((w, (u, v)), shape=(n, n))
    AT = Adj.T.tocsr()
    AT.sort_indices()
    if fmt == "csr":
        A = gg.SparseGPUMatrixCSR(AT, gg.ROCBackend(0))
    else:
        A = gg.SparseGPUMatrixSELL(AT, gg.ROCBackend(0), 64, relabel=True)
    Ao = orc.from_scipy_csr(AT)
    with S.GraphServer(budget_bytes=1 << 20) as srv:
        gid = srv.add_graph(A)
        srv.start()
        print_case("serve negative cycle", fmt)
        bad = srv.submit(gid, S.SSSP, 1)
        good = srv.submit(gid, S.SSSP, 41)
        np.testing.assert_array_equal(srv.fetch(good, gid, S.SSSP), orc.shortest_path(Ao, 41))
        with pytest.raises(gg.NotConvergedError):
            srv.fetch(bad, gid, S.SSSP)
        assert srv.stats()["failed"] == 1
        # sources 13..64 cannot reach the cycle
        srcs = [int(s) for s in np.random.default_rng(9).integers(13, n + 1, 70)]
        rids = [srv.submit(gid, S.SSSP, s) for s in srcs]
        for rid, s in zip(rids, srcs):
            np.testing.assert_array_equal(srv.fetch(rid, gid, S.SSSP), orc.shortest_path(Ao, s),
                                          err_msg=f"source {s}")
        assert srv.stats()["failed"] == 1
        srv.stop()
