"""Every workload of the benchmark configurations at its stated size, on the
GPU through the C ABI, against the oracle or against what the mathematics
fixes.  Mirrors the reference's own checks (test/spmv.jl:10-70 for FP64
products, test/bfs.jl:10-46 for levels, test/shortest_path.jl:31-126 for
distances) at full size:

- CSR SpMV, n = 65,536, exactly 16 distinct uniform columns per row, fp32,
  100 repetitions, each within 1e-5 relative of the FP64 product;
- SELL-C-sigma (C = 64, sigma = 4096) and the degree-relabelled SELL SpMV on
  R-MAT scale 22 within 1e-5 relative of FP64; BFS levels from 64 random
  non-isolated roots (plain pull, direction-optimising, k = 64 batched)
  bit-exact against a textbook queue BFS, one root against the restated
  pull loop, parents checked as BFS trees;
- Bellman-Ford on the 4096 x 4096 grid with 1% long-range edges, 16
  sources, one at a time and batched, by the exact shortest-path
  certificate, one source against the restated Jacobi loop, and
  delta-stepping from the same sources equal to them;
- batched BFS k = 64 and 256 on R-MAT 20 and 22, k = 32 min-plus on
  weighted R-MAT 22, and 512 packed R-MAT graphs of scale 10..18 (the
  persistent launch and the direction-optimising traversal);
- the serving loop over its 8 resident graphs under Poisson arrivals, every
  completed result checked.

The file sorts after the smaller GPU tests so that a failure here does not
hide them under ``pytest -x``."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

SEED = 20261203
INT32_MAX = np.iinfo(np.int32).max
TOL = 1e-5   # fp32 plus-times vs FP64 product, relative (north star)


def _rel_err(got, ref):
    return float(np.max(np.abs(got.astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-30)))


@pytest.fixture(scope="module")
def env():
    import torch
    import gpugraphs as gg
    import oracle as orc
    from gpugraphs import generators as GEN
    assert torch.cuda.is_available()
    orc.lib()
    return gg, torch, orc, GEN, gg.ROCBackend(0)


@pytest.fixture(scope="module")
def rmat22(env):
    gg, torch, orc, GEN, R0 = env
    print_case("generate rmat22")
    H = GEN.rmat(22, 16, 0.57, 0.19, 0.19, seed=SEED, dtype=np.float32, values=GEN.UNIFORM01)
    assert H.m == 1 << 22
    Hb = GEN.HostCSR(H.m, H.n, H.rowptr, H.colval, np.ones(H.nnz, np.bool_), True)
    deg = H.degrees()
    roots = np.random.default_rng(SEED).choice(np.nonzero(deg)[0], 64, replace=False) + 1
    Ao = orc.csr(H.m, H.n, H.rowptr, H.colval, np.ones(H.nnz, np.uint8))
    print_case("queue bfs oracle, 64 roots of rmat22")
    ref = orc.bfs_levels_queue(Ao, roots)
    return H, Hb, roots.astype(np.int32), Ao, ref


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def test_cfg_csr_spmv_uniform_rows_n65536_100_reps(env):
    gg, torch, orc, GEN, R0 = env
    n = 65536
    U = GEN.uniform_rows(n, 16, seed=SEED, dtype=np.float32)
    deg = U.degrees()
    assert U.nnz == 16 * n and np.all(deg == 16)
    S = U.to_scipy()
    assert S.has_canonical_format  # sorted, no duplicates: 16 distinct columns per row
    x = np.random.default_rng(SEED).random(n, dtype=np.float32)
    ref = S.astype(np.float64) @ x.astype(np.float64)
    A = U.to_csr(R0)
    xd = torch.from_numpy(x).cuda()
    y = torch.zeros(n, dtype=torch.float32, device="cuda:0")
    worst = 0.0
    for rep in range(100):
        if rep % 25 == 0:
            print_case("cfg1 csr spmv rep", rep)
        y.fill_(-1.0)
        gg.gpu_spmv_(y, A, xd)
        worst = max(worst, _rel_err(_host(y), ref))
    assert worst <= TOL, worst
    # the restated reference loop meets the same bound on the same input
    yo = np.zeros(n, np.float32)
    orc.spmv(yo, orc.csr(n, n, U.rowptr, U.colval, U.nzval), x)
    assert _rel_err(yo, ref) <= TOL


@pytest.mark.parametrize("layout", ["sigma4096", "relabel_sorted", "csr"])
def test_cfg_rmat22_spmv_vs_fp64(env, rmat22, layout):
    gg, torch, orc, GEN, R0 = env
    H = rmat22[0]
    print_case("rmat22 spmv", layout)
    if layout == "sigma4096":
        A = H.to_sell(R0, 64, 4096, 256, 4096)
    elif layout == "relabel_sorted":
        A = H.to_sell(R0, 64, 1, 1024, 4096, relabel=True, sort_entries=True)
    else:
        A = H.to_csr(R0)
    x = np.random.default_rng(SEED + 1).random(H.n, dtype=np.float32)
    y = torch.zeros(H.m, dtype=torch.float32, device="cuda:0")
    gg.gpu_spmv_(y, A, torch.from_numpy(x).cuda())
    ref = H.to_scipy().astype(np.float64) @ x.astype(np.float64)
    assert _rel_err(_host(y), ref) <= TOL


def test_cfg_rmat22_bfs_64_roots(env, rmat22):
    gg, torch, orc, GEN, R0 = env
    H, Hb, roots, Ao, ref = rmat22
    # the queue BFS agrees with the restated pull loop of bfs.jl at full size
    print_case("rmat22 pull-loop oracle, root", roots[0])
    np.testing.assert_array_equal(orc.bfs_distances(Ao, int(roots[0])), ref[:, 0])
    A = Hb.to_sell(R0, 64, 1, 1024, 65536, relabel=True, sort_entries=True)
    Acsr = Hb.to_csr(R0)
    Asig = Hb.to_sell(R0, 64, 4096, 256, 65536)
    for i, s in enumerate(roots):
        if i % 16 == 0:
            print_case("rmat22 bfs root", i, s)
        np.testing.assert_array_equal(_host(gg.bfs_distances(A, int(s))), ref[:, i])
        np.testing.assert_array_equal(_host(gg.bfs_distances_do(A, int(s), A_out=Acsr)), ref[:, i])
    for i in (0, 31):
        print_case("rmat22 bfs sigma4096 / csr root", i)
        np.testing.assert_array_equal(_host(gg.bfs_distances(Asig, int(roots[i]))), ref[:, i])
        np.testing.assert_array_equal(_host(gg.bfs_distances_do(Acsr, int(roots[i]))), ref[:, i])
    print_case("rmat22 batched bfs k=64")
    D = _host(gg.multi_source_bfs(A, roots))
    np.testing.assert_array_equal(D, ref)


def _check_parents(H, s, par, lev):
    """parents form a BFS tree: the source's parent is 0, unreached vertices
    keep typemax, every other reached v has an in-neighbour parent one level
    up (A_T row v lists the in-neighbours of v)."""
    n = H.m
    assert par[s - 1] == 0
    reach = lev != INT32_MAX
    assert np.array_equal(par == INT32_MAX, ~reach)
    v = np.nonzero(reach)[0]
    v = v[v != s - 1]
    p = par[v].astype(np.int64)
    assert np.all((p >= 1) & (p <= n))
    assert np.array_equal(lev[p - 1], lev[v] - 1)
    # (v, p) is a stored entry of A_T: search p in row v's sorted columns
    keys = np.repeat(np.arange(n, dtype=np.int64), np.diff(H.rowptr.astype(np.int64))) * n + \
        (H.colval.astype(np.int64) - 1)
    q = v.astype(np.int64) * n + (p - 1)
    pos = np.searchsorted(keys, q)
    assert np.all(pos < len(keys)) and np.array_equal(keys[np.minimum(pos, len(keys) - 1)], q)


def test_cfg_rmat22_bfs_parents_are_bfs_trees(env, rmat22):
    gg, torch, orc, GEN, R0 = env
    H, Hb, roots, Ao, ref = rmat22
    Ai = GEN.HostCSR(H.m, H.n, H.rowptr, H.colval, np.ones(H.nnz, np.int32), True)
    for A, name in ((Ai.to_csr(R0), "csr"), (Ai.to_sell(R0, 64, 4096, 256, 4096), "sell")):
        for i in (0, 7):
            print_case("rmat22 parents", name, roots[i])
            par = _host(gg.bfs_parents(A, int(roots[i])))
            _check_parents(H, int(roots[i]), par, ref[:, i])


def test_cfg_grid4096_bellman_ford_16_sources(env):
    gg, torch, orc, GEN, R0 = env
    from gpugraphs.algorithms import shortest_path_
    print_case("generate grid4096")
    Gd = GEN.grid(4096, 4096, 0.01, 1, 1000, seed=SEED, dtype=np.float32)
    assert Gd.m == 4096 * 4096
    w = Gd.nzval
    assert w.min() >= 1 and w.max() <= 1000 and np.all(w == np.round(w))
    srcs = (np.random.default_rng(SEED + 1).choice(Gd.m, 16, replace=False) + 1).astype(np.int32)
    Go = orc.csr(Gd.m, Gd.n, Gd.rowptr, Gd.colval, Gd.nzval)
    A = Gd.to_sell(R0, 64, 1, 0, 4096, relabel=True, sort_entries=True)
    dist = torch.empty(Gd.m, dtype=torch.float32, device="cuda:0")
    plain = np.empty((Gd.m, 16), np.float32, order="F")
    for i, s in enumerate(srcs):
        if i % 4 == 0:
            print_case("grid4096 plain bf source", i, s)
        shortest_path_(A, int(s), dist)
        plain[:, i] = _host(dist)
    assert orc.sssp_certificate(Go, srcs, plain).tolist() == [0] * 16
    print_case("grid4096 pull-loop oracle source", srcs[0])
    np.testing.assert_array_equal(orc.shortest_path(Go, int(srcs[0])), plain[:, 0])
    print_case("grid4096 batched bf k=16")
    D = gg.shortest_path(A, srcs)
    np.testing.assert_array_equal(_host(D), plain)
    del D
    # delta-stepping over the CSR of A (weights differ per direction)
    S = Gd.to_scipy().T.tocsr()
    S.sort_indices()
    A_out = gg.SparseGPUMatrixCSR(S, R0)
    for i, s in enumerate(srcs):
        if i % 4 == 0:
            print_case("grid4096 delta-stepping source", i, s)
        np.testing.assert_array_equal(_host(gg.shortest_path_delta(A_out, int(s))), plain[:, i])
    print_case("grid4096 delta-stepping, 16 sources in one run")
    np.testing.assert_array_equal(_host(gg.shortest_path_delta(A_out, srcs)), plain)


@pytest.mark.parametrize("scale", [20, 22])
def test_cfg_batched_bfs_k64_k256(env, rmat22, scale):
    gg, torch, orc, GEN, R0 = env
    if scale == 22:
        H, Hb = rmat22[0], rmat22[1]
        Ao = rmat22[3]
    else:
        H = GEN.rmat(20, 16, 0.57, 0.19, 0.19, seed=SEED + 20)
        Hb = H
        Ao = orc.csr(H.m, H.n, H.rowptr, H.colval, np.ones(H.nnz, np.uint8))
    roots = (np.random.default_rng(SEED + 256).choice(np.nonzero(Hb.degrees())[0], 256,
                                                      replace=False) + 1).astype(np.int32)
    print_case("queue bfs oracle 256 roots, scale", scale)
    ref = orc.bfs_levels_queue(Ao, roots)
    A = Hb.to_sell(R0, 64, 1, 1024, 65536, relabel=True, sort_entries=True)
    for k in (64, 256):
        print_case("batched bfs scale", scale, "k", k)
        np.testing.assert_array_equal(_host(gg.multi_source_bfs(A, roots[:k])), ref[:, :k])


def test_cfg_batched_sssp_k32_rmat22(env, rmat22):
    gg, torch, orc, GEN, R0 = env
    H, roots = rmat22[0], rmat22[2]
    w = np.random.default_rng(SEED + 1).integers(1, 1001, H.nnz).astype(np.float32)
    Hw = GEN.HostCSR(H.m, H.n, H.rowptr, H.colval, w, True)
    Wo = orc.csr(H.m, H.n, H.rowptr, H.colval, w)
    srcs = roots[:32]
    A = Hw.to_sell(R0, 64, 1, 1024, 4096, relabel=True, sort_entries=True)
    print_case("rmat22 batched sssp k=32")
    D = _host(gg.shortest_path(A, srcs))
    assert orc.sssp_certificate(Wo, srcs, np.asfortranarray(D)).tolist() == [0] * 32
    print_case("rmat22 pull-loop bf oracle source", srcs[0])
    np.testing.assert_array_equal(orc.shortest_path(Wo, int(srcs[0])), D[:, 0])
    print_case("rmat22 plain sssp sources 0, 31")
    for i in (0, 31):
        np.testing.assert_array_equal(_host(gg.shortest_path(A, int(srcs[i]))), D[:, i])


def test_cfg_packed_512_rmat_graphs(env):
    gg, torch, orc, GEN, R0 = env
    rng = np.random.default_rng(SEED + 2)
    scales = rng.integers(10, 19, 512)
    print_case("generate 512 rmat graphs")
    graphs = [GEN.rmat(int(s), 16, 0.57, 0.19, 0.19, seed=SEED + 100 + i)
              for i, s in enumerate(scales)]
    roots = np.array([int(rng.choice(np.nonzero(g.degrees())[0])) + 1 for g in graphs], np.int32)
    P, off = GEN.pack(graphs)
    A = P.to_csr(R0)
    print_case("packed bfs 512 graphs", P.m, P.nnz)
    D, levels = gg.bfs_packed(A, off, roots)
    Dh = _host(D)
    for i, g in enumerate(graphs):
        ref = orc.bfs_levels_queue(orc.csr(g.m, g.n, g.rowptr, g.colval,
                                           np.ones(g.nnz, np.uint8)), [int(roots[i])])[:, 0]
        got = Dh[int(off[i]):int(off[i]) + g.m]
        np.testing.assert_array_equal(got, ref, err_msg=f"packed graph {i} (scale {scales[i]})")
        assert np.all(Dh[int(off[i]) + g.m:int(off[i + 1])] == INT32_MAX)   # padding rows
        assert levels[i] == ref[ref != INT32_MAX].max()
    # the same pack as one direction-optimising traversal over a relabelled
    # SELL (the benchmark's layout): the same arrays as the checked ones
    A_s = gg.SparseGPUMatrixSELL.build_on_device(A, 64, 1, 1024, 65536, True, True)
    print_case("packed bfs 512 graphs, direction-optimising")
    D2, lv2 = gg.bfs_packed_do(A_s, off, roots, A_out=A)
    np.testing.assert_array_equal(_host(D2), Dh)
    np.testing.assert_array_equal(lv2, levels)


def test_cfg_serving_8_graphs_poisson(env, rmat22):
    gg, torch, orc, GEN, R0 = env
    from gpugraphs import serving as S
    H22 = rmat22[1]
    specs = []   # (name, HostCSR pattern, HostCSR weights, bfs matrix, sssp matrix)
    print_case("serving graphs")
    for sc in (16, 18, 20, 22):
        Hs = H22 if sc == 22 else GEN.rmat(sc, 16, 0.57, 0.19, 0.19, seed=SEED + sc)
        Hi = GEN.HostCSR(Hs.m, Hs.n, Hs.rowptr, Hs.colval, np.ones(Hs.nnz, np.bool_), True)
        M = Hi.to_sell(R0, 64, 1, 1024, 65536, relabel=True, sort_entries=True)
        specs.append((f"rmat{sc}", Hi, Hi, M, M))
    for side in (1024, 2048, 4096):
        Hw = GEN.grid(side, side, 0.01, 1, 1000, seed=SEED + side, dtype=np.int32)
        Hp = GEN.HostCSR(Hw.m, Hw.n, Hw.rowptr, Hw.colval, np.ones(Hw.nnz, np.bool_), True)
        specs.append((f"grid{side}", Hp, Hw,
                      Hp.to_sell(R0, 64, 1, 0, 4096, relabel=True, sort_entries=True),
                      Hw.to_sell(R0, 64, 1, 0, 4096, relabel=True, sort_entries=True)))
    Hu = GEN.random_graph(1 << 16, 16, seed=SEED + 7)
    Mu = Hu.to_sell(R0, 64, 1, 1024, 4096, relabel=True, sort_entries=True)
    specs.append(("uniform16", Hu, Hu, Mu, Mu))
    with S.GraphServer(budget_bytes=8 << 30) as srv:
        route, info = [], {}
        for name, Hp, Hw, Mb, Ms in specs:
            gb = srv.add_graph(Mb)
            gs = gb if Ms is Mb else srv.add_graph(Ms)
            route.append((gb, gs, Mb.m, None))
            info[gb] = (name, Hp)
            info[gs] = (name, Hw) if gs != gb else (name, Hp)
        srv.start()
        print_case("serving poisson 2000 q/s for 0.3 s")
        done, wall, results, submitted = S.run_poisson(srv, route, 2000.0, 0.3, seed=SEED,
                                                       keep_every=1)
        srv.stop()
    assert len(done) == len(submitted) and all(c.status == 0 for c in done)
    assert len(results) == len(submitted)
    by = {}
    for rid, g, kind, src in submitted:
        by.setdefault((g, kind), []).append((rid, src))
    checked = 0
    for (g, kind), items in by.items():
        name, Hg = info[g]
        print_case("serving check", name, "bfs" if kind == S.BFS else "sssp", len(items))
        srcs = np.array([s for _, s in items], np.int32)
        if kind == S.BFS:
            Og = orc.csr(Hg.m, Hg.n, Hg.rowptr, Hg.colval, np.ones(Hg.nnz, np.uint8))
            ref = orc.bfs_levels_queue(Og, srcs)
            for q, (rid, _) in enumerate(items):
                np.testing.assert_array_equal(results[rid], ref[:, q],
                                              err_msg=f"served BFS {rid} on {name}")
        else:
            Og = orc.csr(Hg.m, Hg.n, Hg.rowptr, Hg.colval, Hg.nzval)
            Dm = np.empty((Hg.m, len(items)), results[items[0][0]].dtype, order="F")
            for q, (rid, _) in enumerate(items):
                Dm[:, q] = results[rid]
            bad = orc.sssp_certificate(Og, srcs, Dm)
            assert bad.tolist() == [0] * len(items), f"served SSSP on {name}: {bad}"
        checked += len(items)
    assert checked == len(submitted)
