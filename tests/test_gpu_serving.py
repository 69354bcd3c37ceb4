"""GPU tests of the continuous-batching server: every query's result is the
single-source BFS / Bellman-Ford result from its source, whether it runs
alone or in a mix that overflows the engines' slots and the memory budget,
and a query that can never fit the budget fails with BudgetExceeded."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graphs():
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
    mats = [g0.to_csr(R), g1.to_sell(R, 64, 1, 8, 64, relabel=True), g2.to_sell(R, 64, 1, 64, 256),
            g3.to_sell(R, 64, 1, 128, 256, relabel=True)]
    return gg, torch, mats


def _expected(gg, torch, A, kind, src):
    from gpugraphs import serving as S
    out = gg.bfs_distances(A, src) if kind == S.BFS else gg.shortest_path(A, src)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _queries(rng, mats, count):
    from gpugraphs import serving as S
    q = []
    for _ in range(count):
        g = int(rng.integers(len(mats)))
        q.append((g, S.BFS if rng.random() < 0.5 else S.SSSP, int(rng.integers(mats[g].m)) + 1))
    return q


def test_server_results_equal_alone_and_in_mix(graphs):
    gg, torch, mats = graphs
    from gpugraphs import serving as S
    rng = np.random.default_rng(5)
    qs = _queries(rng, mats, 300)          # > 64 BFS slots / 32 SSSP columns per graph
    exp = {i: _expected(gg, torch, mats[g], k, s) for i, (g, k, s) in enumerate(qs)}
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


def test_server_budget_admission_and_rejection(graphs):
    gg, torch, mats = graphs
    from gpugraphs import serving as S
    A, Big = mats[1], mats[2]
    n = A.m
    rng = np.random.default_rng(9)
    srcs = rng.integers(1, n + 1, 40)
    exp = [_expected(gg, torch, A, S.BFS, int(s)) for s in srcs]
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
