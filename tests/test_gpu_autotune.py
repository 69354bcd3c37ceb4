"""The autotuner (the library's gg_tune_layout / gg_tune_push_threshold /
gg_tune_sssp_threshold behind gpugraphs.autotune) times every candidate layout / switch on the
device, keeps the fastest, caches it per graph signature, and the tuned
choice computes the same results as the oracle."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env(oracle):
    import torch
    import gpugraphs as gg
    assert torch.cuda.is_available()
    return gg, torch, oracle


@pytest.mark.parametrize("workload", ["spmv", "bfs", "msbfs", "sssp", "msssp"])
def test_tune_layout_picks_a_timed_candidate(env, workload, tmp_path):
    gg, torch, orc = env
    from gpugraphs import autotune as AT, generators as GEN
    R = GEN.rmat(12, seed=5, dtype=np.float32, values=GEN.UNIFORM01) if workload == "spmv" else \
        GEN.random_graph(3000, 8, seed=6, dtype=np.int32, values=GEN.INTEGER, wmin=1, wmax=99) \
        if workload in ("sssp", "msssp") else GEN.rmat(12, seed=5)
    roots = (np.nonzero(R.degrees())[0][:8] + 1).astype(np.int32)
    cache = str(tmp_path / "tune.json")
    tuner = AT.Autotuner(gg.ROCBackend(0), cache_path=cache, reps=2)
    print_case("autotune", workload)
    res = tuner.tune_layout(R, workload, roots=roots)
    assert res.best in [c for c, _ in res.table]
    assert res.best_ms == min(ms for _, ms in res.table) > 0
    again = AT.Autotuner(gg.ROCBackend(0), cache_path=cache).tune_layout(R, workload, roots=roots)
    assert again.best == res.best                      # served from the cache file
    A = AT.build(R, gg.ROCBackend(0), res.best)
    S = R.to_scipy()
    if workload == "spmv":
        x = np.random.default_rng(1).random(R.n, dtype=np.float32)
        y = torch.zeros(R.m, dtype=torch.float32, device="cuda:0")
        gg.gpu_spmv_(y, A, torch.from_numpy(x).cuda())
        np.testing.assert_allclose(y.cpu().numpy(), S.astype(np.float64) @ x, rtol=1e-5, atol=1e-7)
    elif workload == "sssp":
        np.testing.assert_array_equal(gg.shortest_path(A, int(roots[0])).cpu().numpy(),
                                      orc.shortest_path(orc.from_scipy_csr(S), int(roots[0])))
    elif workload == "msssp":
        np.testing.assert_array_equal(gg.shortest_path(A, roots).cpu().numpy(),
                                      orc.shortest_path_multi(orc.from_scipy_csr(S), roots))
    else:
        np.testing.assert_array_equal(gg.bfs_distances(A, int(roots[0])).cpu().numpy(),
                                      orc.bfs_distances(orc.from_scipy_csr(S), int(roots[0])))


def test_tune_direction_and_graph_steps(env):
    gg, torch, orc = env
    from gpugraphs import autotune as AT, generators as GEN
    R = GEN.rmat(13, seed=8)
    A = R.to_csr(gg.ROCBackend(0))
    roots = (np.nonzero(R.degrees())[0][:4] + 1).astype(np.int32)
    tuner = AT.Autotuner(gg.ROCBackend(0), reps=1)
    res = tuner.tune_direction(A, A, roots, alphas=(4.0, 15.0), betas=(18.0,))
    assert set(res.best) == {"alpha", "beta"} and len(res.table) == 2
    ref = gg.bfs_distances(A, int(roots[0])).cpu().numpy()
    got = gg.bfs_distances_do(A, int(roots[0]), A_out=A, **res.best).cpu().numpy()
    np.testing.assert_array_equal(got, ref)
    from gpugraphs import _lib
    try:
        S = R.to_sell(gg.ROCBackend(0), 64, 1, 256, 1024, relabel=True)
        srcs = (np.nonzero(R.degrees())[0][:40] + 1).astype(np.int32)
        pa = tuner.tune_push_threshold(lambda: gg.multi_source_bfs(S, srcs), alphas=(0.0, 16.0))
        assert pa.best["alpha"] in (0.0, 16.0)
        # the same search in the library (gg_tune_push_threshold)
        pn = tuner.tune_push_threshold(S, alphas=(0.0, 4.0, 16.0), roots=srcs)
        assert pn.best["alpha"] in (0.0, 4.0, 16.0) and len(pn.table) == 3
        assert pn.best_ms == min(ms for _, ms in pn.table) > 0
        D = gg.multi_source_bfs(S, srcs).cpu().numpy()
        np.testing.assert_array_equal(D[:, 0], gg.bfs_distances(A, int(srcs[0])).cpu().numpy())
        gs = tuner.tune_graph_steps(lambda: gg.bfs_distances(A, int(roots[1])), steps=(0, 4))
        assert gs.best["steps"] in (0, 4)
        np.testing.assert_array_equal(gg.bfs_distances(A, int(roots[0])).cpu().numpy(), ref)
    finally:
        _lib.set_graph_steps(_lib.context(0), 0)
        _lib.set_push_threshold(_lib.context(0), 16.0)


def test_tune_sssp_threshold_and_every_switch_matches_oracle(env):
    """gg_tune_sssp_threshold times the k-source Bellman-Ford under each
    sparse-iteration switch and keeps the fastest on the context; whatever
    the switch (0 = every iteration dense, 1e-9 = sparse whenever the fallen
    rows have fewer out-edges than 1e9 times the entries, i.e. always), the
    distances are the oracle's."""
    gg, torch, orc = env
    from gpugraphs import _lib, autotune as AT, generators as GEN
    R = GEN.rmat(12, seed=9)
    w = np.random.default_rng(4).integers(1, 100, R.nnz).astype(np.int32)
    H = GEN.HostCSR(R.m, R.n, R.rowptr, R.colval, w, True)   # symmetric pattern
    S = H.to_sell(gg.ROCBackend(0), 64, 1, 128, 256, relabel=True)
    srcs = (np.nonzero(R.degrees())[0][:24] + 1).astype(np.int32)
    ref = orc.shortest_path_multi(orc.from_scipy_csr(H.to_scipy()), srcs)
    tuner = AT.Autotuner(gg.ROCBackend(0), reps=1)
    try:
        print_case("tune sssp threshold", R.m, len(srcs))
        res = tuner.tune_sssp_threshold(S, srcs, alphas=(0.0, 16.0, 1e-9))
        assert res.best["alpha"] in (0.0, 16.0, 1e-9) and len(res.table) == 3
        assert res.best_ms == min(ms for _, ms in res.table) > 0
        np.testing.assert_array_equal(gg.shortest_path(S, srcs).cpu().numpy(), ref)
        for a in (0.0, 1e-9):
            _lib.set_sssp_threshold(_lib.context(0), a)
            np.testing.assert_array_equal(gg.shortest_path(S, srcs).cpu().numpy(), ref,
                                          err_msg=f"sparse-iteration alpha {a}")
    finally:
        _lib.set_sssp_threshold(_lib.context(0), 16.0)
