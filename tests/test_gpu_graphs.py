"""Captured loops (gg_ctx_set_graph_steps): BFS levels, multi-source BFS and
Bellman-Ford run as HIP graphs relaunched until a device-side convergence
flag is set.  Every replay must give the host-stepped result; graphs are
cached per matrix and workspace, so the same call is replayed several times
with eager work in between, and a changed matrix or size is captured anew."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import print_case
import graphs as G

pytestmark = pytest.mark.gpu


@pytest.fixture
def env(oracle):
    import torch
    import gpugraphs as gg
    from gpugraphs import _lib
    assert torch.cuda.is_available()
    ctx = _lib.context(0)
    yield gg, torch, oracle, ctx, _lib
    _lib.set_graph_steps(ctx, 0)


def host(t):
    return t.cpu().numpy()


def _mats(gg, S):
    R = gg.ROCBackend(0)
    return [("csr", gg.SparseGPUMatrixCSR(S, R)),
            ("sell_ref", gg.SparseGPUMatrixSELL(S, R, 32)),
            ("sell_relabel", gg.SparseGPUMatrixSELL(S, R, 64, heavy_threshold=8,
                                                    heavy_seg_len=16, relabel=True))]


@pytest.mark.parametrize("steps", [2, 7])
def test_graph_bfs_levels_replays(env, steps):
    gg, torch, orc, ctx, L = env
    Adj = G.undirected_adj(2000, G.barabasi_albert(2000, 3, 5), np.bool_)
    AT = Adj.T.tocsr()
    Ao = orc.from_scipy_csr(AT)
    for name, A in _mats(gg, AT):
        for rep in range(3):
            for s in (1, 777, 2000):
                print_case("graph bfs", steps, name, rep, s)
                L.set_graph_steps(ctx, 0)
                eager = host(gg.bfs_distances(A, s))
                L.set_graph_steps(ctx, steps)
                got = host(gg.bfs_distances(A, s))
                np.testing.assert_array_equal(got, eager)
                np.testing.assert_array_equal(got, orc.bfs_distances(Ao, s))


@pytest.mark.parametrize("steps", [2, 8])
def test_graph_multi_source_bfs_replays(env, steps):
    gg, torch, orc, ctx, L = env
    from gpugraphs import generators as GEN
    R = GEN.rmat(12, seed=3)
    S = R.to_scipy()
    Ao = orc.from_scipy_csr(S)
    rng = np.random.default_rng(steps)
    srcs = rng.choice(np.nonzero(R.degrees())[0], 70, replace=False).astype(np.int32) + 1
    for A in (R.to_csr(gg.ROCBackend(0)), R.to_sell(gg.ROCBackend(0), 64, 1, 64, 128, relabel=True)):
        L.set_graph_steps(ctx, 0)
        eager = host(gg.multi_source_bfs(A, srcs))
        L.set_graph_steps(ctx, steps)
        for rep in range(3):
            print_case("graph msbfs", steps, rep)
            np.testing.assert_array_equal(host(gg.multi_source_bfs(A, srcs)), eager)
            np.testing.assert_array_equal(host(gg.bfs_distances(A, int(srcs[rep]))), eager[:, rep])
        for i in (0, 69):
            np.testing.assert_array_equal(eager[:, i], orc.bfs_distances(Ao, int(srcs[i])))


def test_graph_multi_source_bfs_deep_levels(env):
    """Past level 254 the captured byte-level run stops and the batch is
    traversed again with Int32 levels (also captured)."""
    gg, torch, orc, ctx, L = env
    n = 600
    S = sp.diags([np.ones(n - 1), np.ones(n - 1)], [-1, 1], format="csr").astype(np.int32)
    A = gg.SparseGPUMatrixSELL(S, gg.ROCBackend(0), 64, heavy_threshold=8, relabel=True)
    srcs = np.array([1, n, 300], np.int32)
    L.set_graph_steps(ctx, 6)
    D = host(gg.multi_source_bfs(A, srcs))
    Ao = orc.from_scipy_csr(S)
    for i, s in enumerate(srcs):
        np.testing.assert_array_equal(D[:, i], orc.bfs_distances(Ao, int(s)))


@pytest.mark.parametrize("steps", [2, 8])
def test_graph_shortest_path_replays(env, steps):
    gg, torch, orc, ctx, L = env
    from gpugraphs import generators as GEN
    rng = np.random.default_rng(21)
    cases = [G.random_digraph(rng, 400, 0.01, wmax=1000, dtype=np.float32),
             GEN.grid(40, 50, 0.01, 1, 1000, seed=2, dtype=np.int32).to_scipy().T.tocsr().T]
    for Adj in cases:
        AT = sp.csr_matrix(Adj.T)
        Ao = orc.from_scipy_csr(AT)
        srcs = np.array([1, 17, AT.shape[0]], np.int32)
        for name, A in _mats(gg, AT):
            L.set_graph_steps(ctx, 0)
            eager_multi = host(gg.shortest_path(A, srcs))
            for rep in range(2):
                L.set_graph_steps(ctx, steps)
                print_case("graph sssp", steps, name, rep)
                for i, s in enumerate(srcs):
                    d = host(gg.shortest_path(A, int(s)))
                    np.testing.assert_array_equal(d, orc.shortest_path(Ao, int(s)))
                    np.testing.assert_array_equal(d, eager_multi[:, i])
                np.testing.assert_array_equal(host(gg.shortest_path(A, srcs)), eager_multi)
                L.set_graph_steps(ctx, 0)
                np.testing.assert_array_equal(host(gg.shortest_path(A, srcs)), eager_multi)


def test_graph_shortest_path_negative_cycle(env):
    gg, torch, orc, ctx, L = env
    # 1 -> 2 -> 3 -> 1 with total weight -3
    A_T = sp.csr_matrix((np.array([-1, -1, -1], np.int32), (np.array([1, 2, 0]), np.array([0, 1, 2]))),
                        shape=(3, 3))
    A = gg.SparseGPUMatrixCSR(A_T, gg.ROCBackend(0))
    for steps in (0, 4):
        L.set_graph_steps(ctx, steps)
        with pytest.raises(gg.NotConvergedError):
            gg.shortest_path(A, 1)


def test_edge_list_and_matrix_market_inputs(env, tmp_path):
    """Graphs read from files (an undirected weighted edge list, a symmetric
    MatrixMarket pattern) go through the relabelled SELL: BFS levels and
    Bellman-Ford distances equal the oracle's on the same CSR."""
    import scipy.io
    gg, torch, orc, ctx, _lib = env
    from gpugraphs import generators as GEN
    rng = np.random.default_rng(11)
    n, e = 3000, 12000
    u, v, w = rng.integers(1, n + 1, e), rng.integers(1, n + 1, e), rng.integers(1, 100, e)
    p = tmp_path / "g.txt"
    p.write_text("".join(f"{a} {b} {c}\n" for a, b, c in zip(u, v, w)))
    H = GEN.read_edge_list(str(p), base=1, directed=False, dtype=np.int32)
    S = sp.random(n, n, density=0.002, format="coo", random_state=rng)
    q = str(tmp_path / "m.mtx")
    scipy.io.mmwrite(q, sp.triu(S).tocoo(), field="pattern", symmetry="symmetric")
    M = GEN.read_matrix_market(q)
    for name, X in (("edge list", H), ("matrix market", M)):
        assert X.symmetric
        A = X.to_sell(gg.ROCBackend(0), 64, 1, 64, 128, relabel=True)
        Ao = orc.from_scipy_csr(X.to_scipy())
        src = int(np.nonzero(X.degrees())[0][0]) + 1
        print_case("file input", name, X.m, X.nnz)
        np.testing.assert_array_equal(host(gg.bfs_distances(A, src)), orc.bfs_distances(Ao, src))
        if X.nzval.dtype != np.bool_:
            np.testing.assert_array_equal(host(gg.shortest_path(A, src)),
                                          orc.shortest_path(Ao, src))
