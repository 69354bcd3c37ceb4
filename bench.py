#!/usr/bin/env python3
"""BFS queries/s (batched vs plain) and SELL SpMV GB/s on an R-MAT scale-22 graph.

Workload (one GPU): R-MAT scale 22, edge factor 16, a/b/c/d = .57/.19/.19/.05,
symmetrised, self-loops and duplicates removed, ids permuted, Int32 indices;
SELL-C-sigma with C = 64, rows relabelled by decreasing degree (internal ids;
rows longer than the heavy threshold in the segmented partition), or
sigma = 4096 windows without relabelling (--no-relabel).  A step is one batched BFS from 64 random
non-isolated roots (or-and SpMM over 64-bit source masks); the plain
one-root-at-a-time BFS, the direction-optimising BFS and the fp32 plus-times
SELL SpMV are measured beside it, with the oracle on the host cores as the
CPU baseline.  Inputs are resident in HBM before any timed region.

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpugraphs.jl_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs), q))


def sssp_certificate(rowptr, colval, w, d, src):
    """Exact shortest-path certificate at any size: d[src] = 0, no in-edge can
    lower a distance (d[v] <= d[u] + w(u->v) for every edge), and every other
    reached vertex has a tight in-edge (d[v] == min_u d[u] + w(u->v)).  With
    integer-valued weights and sums below 2^24 the float64 check is exact."""
    import numpy as np
    d64 = d.astype(np.float64)
    if np.issubdtype(d.dtype, np.integer):
        d64[d == np.iinfo(d.dtype).max] = np.inf
    cand = d64[colval - 1] + w.astype(np.float64)
    starts = rowptr[:-1].astype(np.int64) - 1
    nz = np.diff(rowptr.astype(np.int64)) > 0
    best = np.full(len(d), np.inf)
    best[nz] = np.minimum.reduceat(cand, starts[nz])
    reached = np.isfinite(d64)
    reached[src - 1] = False
    return bool(d64[src - 1] == 0 and np.all(d64 <= best) and np.all(best[reached] == d64[reached]))


def measured_traffic(args, relabel, kernel):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    runs (profiles/hbm_traffic.json) when this run is the workload they
    measured (the default R-MAT 22 configuration), else None."""
    default = (args.scale == 22 and args.roots == 64 and relabel and args.heavy == 1024 and
               args.bfs_seg == 65536 and args.seg == 4096 and args.push_alpha == 16.0 and
               args.sort_entries and not args.rcm)
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not default or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)[kernel]["traffic_bytes"]


def kernel_pass(_lib, ctx, fn, graph_steps):
    """(summed ms, launches) of the hot kernel over one host-stepped run of fn:
    HIP events on the context stream around each launch (single launches
    inside a captured graph cannot be timed; the kernels are the same)."""
    _lib.set_graph_steps(ctx, 0)
    _lib.enable_timing(ctx, True)
    _lib.hot_kernel_stats(ctx, reset=True)
    fn()
    kms, kn = _lib.hot_kernel_stats(ctx, reset=True)
    _lib.enable_timing(ctx, False)
    _lib.set_graph_steps(ctx, graph_steps)
    return kms, kn


def _time_queries(torch, fn, items):
    lat, outs = [], []
    for it in items:
        t0 = time.perf_counter()
        r = fn(it)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        outs.append(r)
    return lat, outs


def sssp_leg(args, np, torch, gg, G, R0, ctx, _lib, dev, Rg, roots, seed, sigma, relabel):
    """Bellman-Ford one source at a time vs k sources in one multi-source run,
    on the 4-neighbour grid (+1% long-range edges, integer weights 1..1000 as
    fp32) and on the R-MAT graph with the same kind of weights.  The
    k-source run's sparse-iteration switch is tuned per graph first
    (gg_tune_sssp_threshold); both forms are timed warm."""
    from gpugraphs import autotune as AT
    from gpugraphs.algorithms import shortest_path_
    out = {}
    rng = np.random.default_rng(seed + 1)
    t = time.time()
    Gd = G.grid(args.grid, args.grid, 0.01, 1, 1000, seed=seed, dtype=np.float32)
    w_r = rng.integers(1, 1001, Rg.nnz).astype(np.float32)
    Rw = G.HostCSR(Rg.m, Rg.n, Rg.rowptr, Rg.colval, w_r, Rg.symmetric)   # same symmetric pattern
    se = args.sort_entries
    cases = [("grid", Gd, Gd.to_sell(R0, 64, sigma, 0, args.seg, relabel, se),
              (rng.choice(Gd.m, args.sssp_sources, replace=False) + 1).astype(np.int32)),
             ("rmat", Rw, Rw.to_sell(R0, 64, sigma, args.heavy, args.seg, relabel, se),
              roots[:args.sssp_k].astype(np.int32))]
    torch.cuda.synchronize()
    log(f"SSSP graphs: grid{args.grid} n={Gd.m} nnz={Gd.nnz}, weighted rmat; {time.time() - t:.1f}s")
    tuner = AT.Autotuner(R0, reps=2)
    for name, H, A, srcs in cases:
        m, k = H.m, len(srcs)
        sa = tuner.tune_sssp_threshold(A, srcs)
        dist = torch.empty(m, dtype=torch.float32, device=dev)
        shortest_path_(A, int(srcs[0]), dist)
        torch.cuda.synchronize()
        iters = []

        def one(s):
            iters.append(shortest_path_(A, int(s), dist))
            return dist.clone()
        lat, plain = _time_queries(torch, one, srcs)
        kms, kn = kernel_pass(_lib, ctx, lambda: shortest_path_(A, int(srcs[0]), dist),
                              args.graph_steps)
        Dm = torch.empty(k * m, dtype=torch.float32, device=dev).view(k, m).T
        shortest_path_(A, srcs, Dm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it_b = shortest_path_(A, srcs, Dm)
        torch.cuda.synchronize()
        t_b = time.perf_counter() - t0
        same = all(torch.equal(Dm[:, i], plain[i]) for i in range(k))
        exact = sssp_certificate(H.rowptr, H.colval, H.nzval, plain[0].cpu().numpy(), int(srcs[0]))
        assert same and exact, f"SSSP {name}: batched equals plain {same}, certificate {exact}"
        del Dm
        # delta-stepping one source at a time, push over the CSR of A (the
        # transpose of A_T: the weights differ per direction), default delta
        S = H.to_scipy().T.tocsr()
        S.sort_indices()
        A_out = gg.SparseGPUMatrixCSR(S, R0)
        del S
        gg.shortest_path_delta(A_out, int(srcs[0]))
        torch.cuda.synchronize()
        phases = []

        def one_delta(s):
            d, ph = gg.shortest_path_delta(A_out, int(s), return_phases=True)
            phases.append(ph)
            return d
        lat_d, dres = _time_queries(torch, one_delta, srcs)
        same_d = all(torch.equal(dres[i], plain[i]) for i in range(k))
        assert same_d, f"SSSP {name}: delta-stepping differs from Bellman-Ford"
        del dres
        # the k sources as one multi-source delta-stepping run (shared phases)
        gg.shortest_path_delta(A_out, srcs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Dd, ph_k = gg.shortest_path_delta(A_out, srcs, return_phases=True)
        torch.cuda.synchronize()
        t_dk = time.perf_counter() - t0
        same_dk = all(torch.equal(Dd[:, i], plain[i]) for i in range(k))
        assert same_dk, f"SSSP {name}: multi-source delta-stepping differs from Bellman-Ford"
        del Dd, A_out
        B_it = 8 * H.nnz + 4 * (m + 1) + 4 * m + 8 * m  # one full Jacobi iteration
        out[name] = dict(n=m, nnz=H.nnz, sources=k, plain_qps=k / sum(lat),
                         plain_p50_ms=pct(lat, 50) * 1e3, plain_p99_ms=pct(lat, 99) * 1e3,
                         plain_iters_mean=float(np.mean(iters)), iter_kernel_us=kms / max(kn, 1) * 1e3,
                         full_iter_bytes=B_it, batched_qps=k / t_b, batched_ms=t_b * 1e3,
                         batched_iters=it_b, batched_vs_plain=(k / t_b) / (k / sum(lat)),
                         sparse_alpha=sa.best["alpha"],
                         sparse_alpha_ms=[[a, round(t, 3)] for a, t in sa.table],
                         delta_qps=k / sum(lat_d), delta_p50_ms=pct(lat_d, 50) * 1e3,
                         delta_p99_ms=pct(lat_d, 99) * 1e3, delta_phases_mean=float(np.mean(phases)),
                         delta_vs_plain=(k / sum(lat_d)) / (k / sum(lat)),
                         delta_batched_qps=k / t_dk, delta_batched_ms=t_dk * 1e3,
                         delta_batched_phases=ph_k,
                         delta_batched_vs_plain=(k / t_dk) / (k / sum(lat)),
                         exact_certificate=exact)
        log(f"SSSP {name}: plain {k / sum(lat):.1f} q/s (p99 {pct(lat, 99) * 1e3:.1f} ms, "
            f"{np.mean(iters):.0f} iters, {kms / max(kn, 1) * 1e3:.0f} us/iter); batched k={k} "
            f"{k / t_b:.1f} q/s ({it_b} iters) = {(k / t_b) / (k / sum(lat)):.1f}x, "
            f"sparse alpha {sa.best['alpha']}; delta-stepping {k / sum(lat_d):.1f} q/s "
            f"(p99 {pct(lat_d, 99) * 1e3:.1f} ms, {np.mean(phases):.0f} phases) = "
            f"{(k / sum(lat_d)) / (k / sum(lat)):.1f}x one-source Bellman-Ford; {k} sources in one "
            f"delta-stepping run {k / t_dk:.1f} q/s ({ph_k} phases) = "
            f"{(k / t_dk) / (k / sum(lat)):.1f}x")
        del plain
    _lib.set_sssp_threshold(ctx, 16.0)
    return out


def packed_leg(args, np, torch, gg, G, R0, seed):
    """512 independent R-MAT graphs, scale uniform in 10..18 (edge factor 16,
    same generator), packed block-diagonally: one BFS per graph from a random
    non-isolated root, all in one persistent launch, against the same BFS
    run one graph per call on a sample of the graphs."""
    rng = np.random.default_rng(seed + 2)
    t = time.time()
    scales = rng.integers(10, 19, args.packed_graphs)
    graphs = [G.rmat(int(s), 16, 0.57, 0.19, 0.19, seed=seed + 100 + i)
              for i, s in enumerate(scales)]
    roots = np.array([int(rng.choice(np.nonzero(g.degrees())[0])) + 1 for g in graphs], np.int32)
    P, off = G.pack(graphs)
    A = P.to_csr(R0)
    torch.cuda.synchronize()
    log(f"packed {len(graphs)} graphs: {P.m} rows, {P.nnz} entries; {time.time() - t:.1f}s")
    gg.bfs_packed(A, off, roots)
    torch.cuda.synchronize()
    lat = []
    for _ in range(max(args.steps, 1)):
        t0 = time.perf_counter()
        D, levels = gg.bfs_packed(A, off, roots)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
    t_pack = float(np.mean(lat))
    # the same pack as ONE direction-optimising traversal seeded with every
    # graph's source, pulling over a relabelled SELL of the pack (the BFS
    # layout of the single-graph legs) and pushing over the packed CSR
    A_s = gg.SparseGPUMatrixSELL.build_on_device(A, 64, 1, args.heavy, args.bfs_seg, True, True)
    D_do, lv_do = gg.bfs_packed_do(A_s, off, roots, A_out=A)
    torch.cuda.synchronize()
    assert torch.equal(D_do, D) and np.array_equal(lv_do, levels), \
        "direction-optimising packed BFS differs from the persistent launch"
    del D_do
    lat_do = []
    for _ in range(max(args.steps, 1)):
        t0 = time.perf_counter()
        D_do, lv_do = gg.bfs_packed_do(A_s, off, roots, A_out=A)
        torch.cuda.synchronize()
        lat_do.append(time.perf_counter() - t0)
    del D_do, A_s
    t_do = float(np.mean(lat_do))
    sample = rng.choice(len(graphs), min(args.packed_plain, len(graphs)), replace=False)
    mats = [(i, graphs[i].to_csr(R0)) for i in sample]
    gg.bfs_distances(mats[0][1], int(roots[mats[0][0]]))
    torch.cuda.synchronize()
    plain_lat, same = [], True
    for i, Ai in mats:
        t0 = time.perf_counter()
        d = gg.bfs_distances(Ai, int(roots[i]))
        torch.cuda.synchronize()
        plain_lat.append(time.perf_counter() - t0)
        same &= torch.equal(d, D[int(off[i]):int(off[i]) + graphs[i].m])
    assert same, "packed BFS differs from the per-graph BFS"
    q_pack = len(graphs) / t_pack
    q_plain = len(mats) / sum(plain_lat)
    q_do = len(graphs) / t_do
    log(f"packed BFS: {len(graphs)} graphs in {t_pack * 1e3:.2f} ms = {q_pack:.0f} q/s "
        f"(persistent launch), {t_do * 1e3:.2f} ms = {q_do:.0f} q/s (one direction-optimising "
        f"traversal); one graph per call {q_plain:.0f} q/s (p99 {pct(plain_lat, 99) * 1e3:.2f} ms); "
        f"{q_pack / q_plain:.1f}x and {q_do / q_plain:.1f}x; max levels {int(levels.max())}")
    del A, D
    return dict(graphs=len(graphs), rows=int(P.m), entries=int(P.nnz), launch_ms=t_pack * 1e3,
                qps=q_pack, plain_qps=q_plain, plain_p99_ms=pct(plain_lat, 99) * 1e3,
                batched_vs_plain=q_pack / q_plain, do_ms=t_do * 1e3, do_qps=q_do,
                do_p99_ms=pct(lat_do, 99) * 1e3, do_vs_plain=q_do / q_plain,
                max_levels=int(levels.max()), plain_sample=len(mats))


def tuned_leg(args, np, torch, gg, G, R0, ctx, _lib, seed):
    """The batched forms on R-MAT scale 20 (same generator) with the layout
    and the push/pull threshold picked by the autotuner: its multi-source BFS
    candidates timed on 64 roots, then the threshold; then batched BFS
    (k = 64 and 256) against the plain BFS, and k = 32 Bellman-Ford
    (integer weights 1..1000 as fp32) against one source at a time, each on
    its tuned layout, the k-source run with its tuned sparse-iteration
    switch."""
    from gpugraphs import autotune as AT
    from gpugraphs.algorithms import shortest_path_
    sc = args.tuned_scale
    t = time.time()
    H = G.rmat(sc, 16, 0.57, 0.19, 0.19, seed=seed + 20)
    Hb = G.HostCSR(H.m, H.n, H.rowptr, H.colval, np.ones(H.nnz, np.bool_), H.symmetric)
    rng = np.random.default_rng(seed + 20)
    nzr = np.nonzero(H.degrees())[0]
    roots = rng.choice(nzr, 256, replace=False).astype(np.int32) + 1
    tuner = AT.Autotuner(R0, reps=2)
    lay = tuner.tune_layout(Hb, "msbfs", roots=roots[:64])
    A = AT.build(Hb, R0, lay.best)
    pa = tuner.tune_push_threshold(A, alphas=(0.0, 4.0, 16.0, 64.0), roots=roots[:64])
    tune_s = time.time() - t
    out = dict(scale=sc, n=int(H.m), nnz=int(H.nnz), tune_s=round(tune_s, 1),
               layout=lay.best, layout_ms=round(lay.best_ms, 3), push_alpha=pa.best["alpha"],
               candidates=len(lay.table))
    for k in (64, 256):
        r = roots[:k]
        gg.multi_source_bfs(A, r)
        torch.cuda.synchronize()
        lat = []
        for _ in range(max(args.steps, 2)):
            t0 = time.perf_counter()
            D = gg.multi_source_bfs(A, r)
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t0)
        out[f"msbfs_k{k}_qps"] = round(k / float(np.mean(lat)), 1)
        out[f"msbfs_k{k}_p99_ms"] = round(pct(lat, 99) * 1e3, 3)
    plat, same = [], True
    for i in range(16):
        t0 = time.perf_counter()
        d = gg.bfs_distances(A, int(roots[i]))
        torch.cuda.synchronize()
        plat.append(time.perf_counter() - t0)
        same &= bool(torch.equal(d, D[:, i]))
    assert same, "tuned batched BFS differs from the plain BFS"
    out["plain_qps"] = round(16 / sum(plat), 1)
    out["plain_p99_ms"] = round(pct(plat, 99) * 1e3, 3)
    out["batched_vs_plain_k64"] = round(out["msbfs_k64_qps"] / out["plain_qps"], 2)
    # k = 32 Bellman-Ford on the layout the tuner picks for Bellman-Ford
    w = rng.integers(1, 1001, H.nnz).astype(np.float32)
    Hw = G.HostCSR(H.m, H.n, H.rowptr, H.colval, w, H.symmetric)
    srcs = roots[:args.sssp_k]
    lw = tuner.tune_layout(Hw, "sssp", roots=srcs)       # one source at a time
    lm = tuner.tune_layout(Hw, "msssp", roots=srcs)      # the k sources at once
    out["sssp_layout"], out["msssp_layout"] = lw.best, lm.best
    Aw = AT.build(Hw, R0, lw.best)
    Am = Aw if lm.best == lw.best else AT.build(Hw, R0, lm.best)
    sa = tuner.tune_sssp_threshold(Am, srcs)                # sparse-iteration switch
    out["sssp_alpha"] = sa.best["alpha"]
    dist = torch.empty(H.m, dtype=torch.float32, device=f"cuda:{R0.device}")
    shortest_path_(Aw, int(srcs[0]), dist)
    torch.cuda.synchronize()
    slat, plain = [], []
    for s_ in srcs:
        t0 = time.perf_counter()
        shortest_path_(Aw, int(s_), dist)
        torch.cuda.synchronize()
        slat.append(time.perf_counter() - t0)
        plain.append(dist.clone())
    k = len(srcs)
    Dm = torch.empty(k * H.m, dtype=torch.float32, device=dist.device).view(k, H.m).T
    shortest_path_(Am, srcs, Dm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    shortest_path_(Am, srcs, Dm)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    assert all(torch.equal(Dm[:, i], plain[i]) for i in range(k)), "tuned batched SSSP differs"
    out["sssp_k"] = k
    out["sssp_plain_qps"] = round(k / sum(slat), 1)
    out["sssp_batched_qps"] = round(k / tb, 1)
    out["sssp_batched_vs_plain"] = round((k / tb) / (k / sum(slat)), 2)
    log(f"tuned R-MAT {sc}: layout {lay.best} ({len(lay.table)} candidates, push alpha "
        f"{pa.best['alpha']}, {tune_s:.1f}s); batched BFS k=64 {out['msbfs_k64_qps']:.0f} q/s, "
        f"k=256 {out['msbfs_k256_qps']:.0f} q/s, plain {out['plain_qps']:.0f} q/s; SSSP k={k} "
        f"{out['sssp_batched_qps']:.0f} q/s vs {out['sssp_plain_qps']:.0f} q/s")
    _lib.set_push_threshold(ctx, args.push_alpha)
    _lib.set_sssp_threshold(ctx, 16.0)
    return out


def serving_leg(args, np, torch, gg, G, R0, seed, sigma, relabel):
    """Continuous batching over 8 HBM-resident graphs: R-MAT scales 16, 18, 20,
    22 (edge factor 16), 4-neighbour grids 1024^2, 2048^2, 4096^2 (+1%
    long-range edges, integer weights 1..1000) and a 2^16-vertex uniform
    random graph of degree 16.  Poisson arrivals; each query is a BFS or an
    SSSP from a uniform random source on a uniform random graph (BFS on a grid
    follows its pattern, SSSP its weights).  Reports q/s and p50/p99 latency
    from arrival to completion, and checks a sample of results against the
    same query run alone."""
    from gpugraphs import serving as S
    t = time.time()
    specs = []   # (name, bfs matrix, sssp matrix)
    for sc in (16, 18, 20, 22):
        H = G.rmat(sc, 16, 0.57, 0.19, 0.19, seed=seed + sc)
        M = H.to_sell(R0, 64, sigma, args.heavy, args.bfs_seg, relabel, args.sort_entries)
        specs.append((f"rmat{sc}", M, M))
    for side in (1024, 2048, 4096):
        H = G.grid(side, side, 0.01, 1, 1000, seed=seed + side, dtype=np.int32)
        Hb = G.HostCSR(H.m, H.n, H.rowptr, H.colval, np.ones(H.nnz, np.bool_), H.symmetric)
        specs.append((f"grid{side}", Hb.to_sell(R0, 64, sigma, 0, args.seg, relabel,
                                                args.sort_entries),
                      H.to_sell(R0, 64, sigma, 0, args.seg, relabel, args.sort_entries)))
    H = G.random_graph(1 << 16, 16, seed=seed + 7)
    M = H.to_sell(R0, 64, sigma, args.heavy, args.seg, relabel, args.sort_entries)
    specs.append(("uniform16", M, M))
    torch.cuda.synchronize()
    log(f"serving graphs built in {time.time() - t:.1f}s")
    budget = int(args.serve_budget_gib * (1 << 30))
    runs = []
    rates = [float(r) for r in str(args.serve_rate).split(",") if r]
    secs = [float(x) for x in str(args.serve_seconds).split(",") if x]
    secs = secs + [secs[-1]] * (len(rates) - len(secs))   # one window per rate
    for rate, window in zip(rates, secs):
        srv = S.GraphServer(budget_bytes=budget)
        srv.set_pacing(args.serve_pacing)
        table, route = [], []
        for name, Mb, Ms in specs:
            gb = srv.add_graph(Mb)
            gs = gb if Ms is Mb else srv.add_graph(Ms)
            route.append((gb, gs, Mb.m, None))
            table.append((name, gb, gs, Mb, Ms))
        srv.start()
        done, wall, results, submitted = S.run_poisson(srv, route, rate, window,
                                                       seed=seed, keep_every=97)
        st = srv.stats()
        srv.stop()
        # per graph and engine: steps, mean busy slots per step, queries
        # finished, device ms per step and per finished query
        engines = {}
        for name, gb, gs, Mb, Ms in table:
            for g, kind, label in ((gb, S.BFS, "bfs"), (gs, S.SSSP, "sssp")):
                es = srv.engine_stats(g, kind)
                if es["steps"]:
                    engines[f"{name}/{label}"] = [
                        es["steps"], round(es["busy_slot_steps"] / es["steps"], 1), es["finished"],
                        round(1e3 * es["device_s"] / es["steps"], 3),
                        round(1e3 * es["device_s"] / max(1, es["finished"]), 3)]
        ok = [c for c in done if c.status == 0]
        lat = [c.latency_s for c in ok]
        by_id = {r: (g, k, s) for r, g, k, s in submitted}
        mats = {}
        for name, gb, gs, Mb, Ms in table:
            mats[gb], mats[gs] = Mb, Ms
        checked = 0
        for rid, out in list(results.items())[:24]:
            g, k, src = by_id[rid]
            ref = gg.bfs_distances(mats[g], src) if k == S.BFS else gg.shortest_path(mats[g], src)
            assert np.array_equal(out, ref.cpu().numpy()), f"served query {rid} differs from alone"
            checked += 1
        srv.close()
        per_kind = {}
        for kind, label in ((S.BFS, "bfs"), (S.SSSP, "sssp")):
            kl = [c.latency_s for c in ok if c.kind == kind]
            if kl:
                per_kind[label] = dict(n=len(kl), p50_ms=round(pct(kl, 50) * 1e3, 2),
                                       p99_ms=round(pct(kl, 99) * 1e3, 2))
        # per graph and kind: count, p50 / p99 latency and mean time from admission to completion
        gname = {}
        for name, gb, gs, Mb, Ms in table:
            gname[gb], gname[gs] = name, name
        per_graph = {}
        for name in dict.fromkeys(gname.values()):
            for kind, label in ((S.BFS, "bfs"), (S.SSSP, "sssp")):
                cs = [c for c in ok if gname[c.graph] == name and c.kind == kind]
                if cs:
                    kl = [c.latency_s for c in cs]
                    per_graph[f"{name}/{label}"] = [len(cs), round(pct(kl, 50) * 1e3, 1),
                                                   round(pct(kl, 99) * 1e3, 1),
                                                   round(float(np.mean([c.complete_s - c.admit_s
                                                                        for c in cs])) * 1e3, 1)]
        run = dict(rate_qps=rate, pacing=args.serve_pacing, arrival_s=window,
                   submitted=len(submitted),
                   completed=len(ok), failed=len(done) - len(ok), wall_s=round(wall, 3),
                   qps=round(len(ok) / wall, 1),
                   p50_ms=round(pct(lat, 50) * 1e3, 2) if lat else None,
                   p99_ms=round(pct(lat, 99) * 1e3, 2) if lat else None, by_kind=per_kind,
                   by_graph_n_p50_p99_service_ms=per_graph,
                   engines_steps_busy_finished_ms_per_step_ms_per_query=engines,
                   device_s_all_engines=round(sum(v[0] * v[3] for v in engines.values()) / 1e3, 3),
                   steps=st["steps"],
                   peak_arena_gib=round(srv.last_run_peaks["arena_used"] / (1 << 30), 2),
                   peak_waiting=srv.last_run_peaks["waiting"],
                   checked_alone=checked)
        runs.append(run)
        log(f"serving at {rate:.0f} q/s offered: {len(ok)}/{len(submitted)} queries in {wall:.1f}s "
            f"= {len(ok) / wall:.0f} q/s, p50 {run['p50_ms']:.1f} ms p99 {run['p99_ms']:.1f} ms, "
            f"peak arena {run['peak_arena_gib']} GiB, peak waiting {run['peak_waiting']}, "
            f"{checked} checked against alone")
    return dict(graphs=[s[0] for s in specs], budget_bytes=budget, runs=runs)


def copy_calibration(torch, dev, nbytes=1 << 30, reps=10):
    """Plain device copy of a known size: FETCH_SIZE / WRITE_SIZE per dispatch of
    this kernel divided by nbytes is the counters' scale factor."""
    src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    log(f"copy {nbytes} B: {ms * 1e3:.1f} us, {2 * nbytes / ms / 1e6:.0f} GB/s (read + write)")
    print(json.dumps({"copy_bytes": nbytes, "us": round(ms * 1e3, 1),
                      "gbs_read_plus_write": round(2 * nbytes / ms / 1e6, 1)}), flush=True)


def cpu_leg(np, G, Rg, roots, seed, gpu_levels):
    """The oracle (the reference's CPU loops restated in C++/OpenMP) on this
    host, on bounded samples of the benchmark workloads: one BFS (the pull
    loop of bfs.jl) and one fp32 SpMV on R-MAT 22 with all the job's threads
    and with one; the CSR SpMV of the n = 65,536 uniform-rows matrix x100
    (all threads and one); one Bellman-Ford (Jacobi loop of
    shortest_path.jl) from one source of the 4096^2 grid; one iteration of
    the k = 32 multi-source Bellman-Ford on weighted R-MAT 22, scaled by the
    iteration count the GPU run needs.  The GPU's BFS levels from the first
    root are checked against the oracle's here (full-size parity)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes as C
    import oracle as orc
    T, nproc = cpu_threads()
    n, nnz = Rg.m, Rg.nnz
    out = dict(threads=T, nproc=nproc)
    Ao = orc.csr(n, n, Rg.rowptr, Rg.colval, np.ones(nnz, np.uint8))
    root = int(roots[0])
    for th, key in ((T, "bfs_s"), (1, "bfs_1t_s")):
        t0 = time.perf_counter()
        dref = orc.bfs_distances(Ao, root, nthreads=th)
        out[key] = time.perf_counter() - t0
    if gpu_levels is not None and gpu_levels[0] == root:
        assert np.array_equal(dref, gpu_levels[1]), "GPU BFS levels differ from the oracle at full size"
    Af = orc.csr(n, n, Rg.rowptr, Rg.colval, Rg.nzval)
    xh = np.random.default_rng(1).random(n, dtype=np.float32)
    yh = np.zeros(n, np.float32)
    B = nnz * 8 + 12 * n
    for th, key in ((T, "spmv_s"), (1, "spmv_1t_s")):
        t0 = time.perf_counter()
        orc.spmv(yh, Af, xh, nthreads=th)
        out[key] = time.perf_counter() - t0
    out["spmv_gbs"] = B / out["spmv_s"] / 1e9
    U = G.uniform_rows(65536, 16, seed=seed, dtype=np.float32)
    Uo = orc.csr(U.m, U.n, U.rowptr, U.colval, U.nzval)
    xu = np.random.default_rng(2).random(U.n, dtype=np.float32)
    yu = np.zeros(U.m, np.float32)
    for th, key in ((T, "csr65536_x100_s"), (1, "csr65536_x100_1t_s")):
        t0 = time.perf_counter()
        for _ in range(100):
            orc.spmv(yu, Uo, xu, nthreads=th)
        out[key] = time.perf_counter() - t0
    Gd = G.grid(4096, 4096, 0.01, 1, 1000, seed=seed, dtype=np.float32)
    Go = orc.csr(Gd.m, Gd.n, Gd.rowptr, Gd.colval, Gd.nzval)
    src = int(np.random.default_rng(seed + 1).integers(Gd.m)) + 1
    it = C.c_int64()
    dist = np.empty(Gd.m, np.float32)
    t0 = time.perf_counter()
    st = orc.lib().orc_shortest_path(C.byref(Go.view()), src, dist.ctypes.data, orc.FLOAT32, T, 0,
                                     C.byref(it))
    out["grid_bf_s"] = time.perf_counter() - t0
    out["grid_bf_iters"] = int(it.value)
    assert st == 0
    del Gd, Go, dist
    w = np.random.default_rng(seed + 1).integers(1, 1001, nnz).astype(np.float32)
    Wo = orc.csr(n, n, Rg.rowptr, Rg.colval, w)
    srcs = np.ascontiguousarray(roots[:32], np.int32)
    D = np.empty((n, 32), np.float32, order="F")
    t0 = time.perf_counter()
    st = orc.lib().orc_shortest_path_multi(C.byref(Wo.view()), srcs.ctypes.data, 32, D.ctypes.data,
                                           orc.FLOAT32, T, 1, C.byref(it))
    out["rmat_bf32_iter_s"] = time.perf_counter() - t0
    assert st in (0, 8)   # 8: stopped at the one-iteration sample
    del D, Wo, w
    log(f"CPU oracle ({T} threads of {nproc}): BFS {out['bfs_s']:.2f} s (1 thread {out['bfs_1t_s']:.2f} s), "
        f"SpMV {out['spmv_s'] * 1e3:.0f} ms (1 thread {out['spmv_1t_s'] * 1e3:.0f} ms), "
        f"CSR n=65536 x100 {out['csr65536_x100_s'] * 1e3:.0f} ms (1 thread "
        f"{out['csr65536_x100_1t_s'] * 1e3:.0f} ms), grid Bellman-Ford {out['grid_bf_s']:.1f} s "
        f"({out['grid_bf_iters']} iterations), k=32 Bellman-Ford {out['rmat_bf32_iter_s']:.1f} s per iteration")
    return out


def hbm_ceiling(torch, dev, nbytes=4 << 30, reps=10):
    """What this device's HBM delivers to a streaming kernel of this project
    (libgg_calibrate.so: 16-byte non-temporal loads / stores on 8 blocks of
    256 threads per CU): GB/s for a pure read, a pure write and a copy
    (read + write bytes) of nbytes, each the mean of reps launches timed with
    HIP events on the current stream."""
    import ctypes as C
    lib = C.CDLL(os.path.join(ROOT, "gpugraphs.jl_amd", "lib", "libgg_calibrate.so"))
    lib.ggc_sink_bytes.restype = C.c_size_t
    for f in (lib.ggc_stream_read, lib.ggc_stream_write, lib.ggc_stream_copy):
        f.restype = C.c_int
    buf = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(buf)
    sink = torch.empty(lib.ggc_sink_bytes() // 4, dtype=torch.float32, device=dev)
    st = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    ms = C.c_double()
    out = {"bytes": nbytes}
    assert lib.ggc_stream_read(st, C.c_void_p(buf.data_ptr()), C.c_size_t(nbytes),
                               C.c_void_p(sink.data_ptr()), reps, C.byref(ms)) == 0
    out["read_gbs"] = round(nbytes / ms.value / 1e6, 1)
    assert lib.ggc_stream_write(st, C.c_void_p(dst.data_ptr()), C.c_size_t(nbytes), reps,
                                C.byref(ms)) == 0
    out["write_gbs"] = round(nbytes / ms.value / 1e6, 1)
    assert lib.ggc_stream_copy(st, C.c_void_p(buf.data_ptr()), C.c_void_p(dst.data_ptr()),
                               C.c_size_t(nbytes), reps, C.byref(ms)) == 0
    out["copy_gbs_read_plus_write"] = round(2 * nbytes / ms.value / 1e6, 1)
    del buf, dst, sink
    log(f"HBM ceiling ({nbytes >> 20} MiB streams): read {out['read_gbs']:.0f} GB/s, write "
        f"{out['write_gbs']:.0f} GB/s, copy {out['copy_gbs_read_plus_write']:.0f} GB/s (read + write)")
    return out


def gather_only_ms(torch, dev, A, reps, wide=False):
    """The SpMV of a SELL matrix without its arithmetic, values or result: its
    column streams (slices, then heavy rows) read in storage order with the
    x gathers they name (libgg_calibrate.so ggc_gather), in ms; wide: 8-byte
    gathers, one dense pass of a multi-source BFS level over 64-source masks
    (ggc_gather8)."""
    import ctypes as C
    lib = C.CDLL(os.path.join(ROOT, "gpugraphs.jl_amd", "lib", "libgg_calibrate.so"))
    lib.ggc_sink_bytes.restype = C.c_size_t
    fn = lib.ggc_gather8 if wide else lib.ggc_gather
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int,
                   C.POINTER(C.c_double)]
    n_int = A.nslices * A.slice_size + A.n_heavy
    xg = (torch.randint(0, 1 << 62, (max(n_int, A.n),), dtype=torch.int64, device=dev) if wide
          else torch.rand(max(n_int, A.n), dtype=torch.float32, device=dev))
    sink = torch.empty(lib.ggc_sink_bytes() // 4, dtype=torch.float32, device=dev)
    st = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    total = 0.0
    for cols in (A.colval, A.heavy_col if A.n_heavy else None):
        if cols is None or cols.numel() == 0:
            continue
        ms = C.c_double()
        assert fn(st, cols.data_ptr(), cols.numel(), xg.data_ptr(), sink.data_ptr(), reps,
                  C.byref(ms)) == 0
        total += ms.value
    return total


def cpu_threads():
    """Threads of the CPU baseline: the job's CPU share (OMP_NUM_THREADS when the
    environment sets it), else every core nproc reports."""
    n = os.cpu_count() or 1
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    env = os.environ.get("OMP_NUM_THREADS", "")
    return (int(env) if env.isdigit() and int(env) > 0 else n), n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--roots", type=int, default=64)
    ap.add_argument("--heavy", type=int, default=None,
                    help="SELL heavy-row threshold (default 1024 relabelled, 256 sigma windows)")
    ap.add_argument("--no-relabel", action="store_true",
                    help="sigma = 4096 windows in original ids instead of degree relabelling")
    ap.add_argument("--seg", type=int, default=4096,
                    help="heavy segment length of the SpMV and Bellman-Ford matrices")
    ap.add_argument("--bfs-seg", type=int, default=65536,
                    help="heavy segment length of the BFS (Bool pattern) matrix")
    ap.add_argument("--plain-roots", type=int, default=64)
    ap.add_argument("--wide-roots", type=int, default=256,
                    help="sources of the second batched BFS measurement (0: skip)")
    ap.add_argument("--push-alpha", type=float, default=16.0,
                    help="multi-source BFS levels run top-down when frontier edges * alpha < entries")
    ap.add_argument("--graph-steps", type=int, default=8,
                    help="levels / iterations per captured graph launch (0: host-stepped loops)")
    ap.add_argument("--spmv-reps", type=int, default=20)
    ap.add_argument("--spmm-k", type=int, default=32, help="columns of the SpMM leg (0: skip)")
    ap.add_argument("--no-sort-entries", dest="sort_entries", action="store_false",
                    help="keep each row's entries in input order (default: by increasing stored "
                         "column: light rows by degree, heavy rows last)")
    ap.add_argument("--rcm", action="store_true",
                    help="relabelled R-MAT layouts put equal-degree rows in reverse Cuthill-McKee "
                         "order (GG_SELL_RCM; default: row order)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--grid", type=int, default=4096, help="Bellman-Ford grid side")
    ap.add_argument("--sssp-sources", type=int, default=16, help="grid Bellman-Ford sources")
    ap.add_argument("--sssp-k", type=int, default=32, help="batched SSSP sources on R-MAT")
    ap.add_argument("--packed-graphs", type=int, default=512,
                    help="R-MAT graphs (scale 10..18) packed into one persistent BFS launch")
    ap.add_argument("--packed-plain", type=int, default=64,
                    help="graphs of the pack also run one BFS call at a time (plain baseline)")
    ap.add_argument("--tuned-scale", type=int, default=20,
                    help="R-MAT scale of the autotuned batched leg (0: skip)")
    ap.add_argument("--serve-rate", type=str, default="1500,2000,10000",
                    help="Poisson arrivals per second, one serving run per comma-separated rate")
    ap.add_argument("--serve-seconds", type=str, default="8,8,1.5",
                    help="arrival window of each serving rate (comma list; the last one repeats)")
    ap.add_argument("--serve-pacing", type=float, default=24.0,
                    help="server step pacing factor (gg_server_set_pacing; 0: step whenever a slot is busy)")
    ap.add_argument("--serve-budget-gib", type=float, default=64.0,
                    help="HBM result-arena budget of the server (GiB)")
    ap.add_argument("--only", default="",
                    help="one part: spmv | spmm | msbfs | bfs | sssp | packed | tuned | serve | cpu, "
                         "copy (a 1 GiB device copy that calibrates the HBM byte counters) or "
                         "stream (the project's streaming read / write / copy kernels)")
    args = ap.parse_args()
    relabel = not args.no_relabel
    sigma = 1 if relabel else 4096
    if args.heavy is None:
        args.heavy = 1024 if relabel else 256
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build "
              "(the per-GPU shard and how it would shard are described in DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import gpugraphs as gg
    from gpugraphs import _lib
    from gpugraphs import generators as G

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.only == "copy":
        copy_calibration(torch, dev)
        return
    if args.only == "stream":
        print(json.dumps(hbm_ceiling(torch, dev, reps=args.spmv_reps)), flush=True)
        return
    R0 = gg.ROCBackend(0)
    seed = 20261203
    # ------------------------------------------------------------- serving
    # First, as a serving process would: the runtime maps a process's HIP
    # streams onto its 4 hardware queues by the order they were created, and
    # after the other legs' streams the engines' streams shared queues less
    # well (1,500 q/s offered: p99 4.2 s against 0.9 s).
    serve_res = None
    if args.only in ("", "serve") and max(float(x) for x in str(args.serve_seconds).split(",") if x) > 0:
        serve_res = serving_leg(args, np, torch, gg, G, R0, seed, sigma, relabel)
        torch.cuda.empty_cache()
    ceiling = hbm_ceiling(torch, dev) if args.only in ("", "spmv") else None

    t = time.time()
    Rg = G.rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=seed, dtype=np.float32, values=G.UNIFORM01)
    n, nnz = Rg.m, Rg.nnz
    deg = Rg.degrees()
    log(f"rmat{args.scale}: n={n} nnz={nnz} max_deg={deg.max()} gen {time.time() - t:.1f}s")

    t = time.time()
    rcm = relabel and args.rcm   # equal-degree rows in reverse Cuthill-McKee order
    A_f = Rg.to_sell(R0, 64, sigma, args.heavy, args.seg, relabel,   # fp32 values for SpMV
                     sort_entries=args.sort_entries, rcm=rcm)
    Rb = G.HostCSR(n, n, Rg.rowptr, Rg.colval, np.ones(nnz, np.bool_), Rg.symmetric)
    A_b = Rb.to_sell(R0, 64, sigma, args.heavy, args.bfs_seg, relabel,   # Bool pattern for BFS
                     sort_entries=args.sort_entries, rcm=rcm)
    A_csr = Rb.to_csr(R0)                                          # out-edges for push (symmetric)
    torch.cuda.synchronize()
    host_build_s = time.time() - t
    log(f"SELL builds + upload {host_build_s:.1f}s; stored/nnz {(A_f.n_stored + (int(A_f.heavy_ptr[-1].item()) if A_f.n_heavy else 0)) / nnz:.3f}")
    # the same BFS layout built on the device from the CSR already in HBM
    gg.SparseGPUMatrixSELL.build_on_device(A_csr, 64, sigma, args.heavy, args.bfs_seg, relabel,
                                           args.sort_entries)
    torch.cuda.synchronize()
    t = time.perf_counter()
    A_d = gg.SparseGPUMatrixSELL.build_on_device(A_csr, 64, sigma, args.heavy, args.bfs_seg,
                                                 relabel, args.sort_entries)
    torch.cuda.synchronize()
    dev_build_ms = (time.perf_counter() - t) * 1e3
    # the device builder has no RCM ties: compare it with the host layout without them
    A_h = Rb.to_sell(R0, 64, sigma, args.heavy, args.bfs_seg, relabel,
                     sort_entries=args.sort_entries) if rcm else A_b
    assert torch.equal(A_d.slice_ptr, A_h.slice_ptr) and torch.equal(A_d.colval, A_h.colval), \
        "device-built SELL differs from the host-built one"
    del A_d, A_h
    log(f"SELL build on the device from the resident CSR: {dev_build_ms:.1f} ms")

    rng = np.random.default_rng(seed)
    roots = rng.choice(np.nonzero(deg)[0], args.roots, replace=False).astype(np.int32) + 1
    ctx = _lib.context(0)
    _lib.set_graph_steps(ctx, args.graph_steps)
    _lib.set_push_threshold(ctx, args.push_alpha)
    res = {"sell_build": dict(host_two_layouts_s=round(host_build_s, 3),
                              device_bfs_layout_ms=round(dev_build_ms, 2))}

    # ---------------------------------------------------------------- SpMV
    if args.only in ("", "spmv"):
        x = torch.from_numpy(rng.random(n, dtype=np.float32)).to(dev)
        y = torch.zeros(n, dtype=torch.float32, device=dev)
        S64 = Rg.to_scipy().astype(np.float64)
        ref = S64 @ x.cpu().numpy().astype(np.float64)
        del S64
        B_spmv = nnz * 8 + 4 * (n + 1) + 4 * n + 4 * n  # algorithmic bytes per SpMV

        def spmv_leg(A):
            """(ms per call, ms per hot-kernel launch, max rel err vs FP64)."""
            for _ in range(3):
                gg.gpu_spmv_(y, A, x)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(args.spmv_reps):
                gg.gpu_spmv_(y, A, x)
            e1.record()
            torch.cuda.synchronize()
            call_ms = e0.elapsed_time(e1) / args.spmv_reps

            def reps():
                for _ in range(args.spmv_reps):
                    gg.gpu_spmv_(y, A, x)
            kms, kn = kernel_pass(_lib, ctx, reps, args.graph_steps)
            err = float(np.max(np.abs(y.cpu().numpy() - ref) / np.maximum(np.abs(ref), 1e-30)))
            assert err <= 1e-5, f"SpMV parity failed: max rel err {err}"
            return call_ms, kms / max(kn, 1), err

        spmv_ms, spmv_kernel_ms, relerr = spmv_leg(A_f)
        res["spmv"] = dict(ms=spmv_ms, kernel_ms=spmv_kernel_ms, bytes=B_spmv,
                           gbs=B_spmv / spmv_ms / 1e6, kernel_gbs=B_spmv / spmv_kernel_ms / 1e6,
                           max_rel_err_vs_fp64=relerr)
        res["spmv"]["gather_only_ms"] = gather_only_ms(torch, dev, A_f, args.spmv_reps)
        log(f"the same column streams with their x gathers alone (no values, no y): "
            f"{res['spmv']['gather_only_ms'] * 1e3:.1f} us")
        log(f"SELL SpMV: {spmv_ms * 1e3:.1f} us/SpMV ({B_spmv / spmv_ms / 1e6:.0f} GB/s), kernel "
            f"{spmv_kernel_ms * 1e3:.1f} us ({B_spmv / spmv_kernel_ms / 1e6:.0f} GB/s), rel err {relerr:.1e}")
        if relabel:
            # the configuration's named layout: C = 64, sigma = 4096 windows (no
            # relabelling), rows longer than 256 entries in the heavy partition
            A_s = Rg.to_sell(R0, 64, 4096, 256, args.seg)
            s_ms, s_kms, s_err = spmv_leg(A_s)
            stored = A_s.n_stored + (int(A_s.heavy_ptr[-1].item()) if A_s.n_heavy else 0)
            s_gather = gather_only_ms(torch, dev, A_s, args.spmv_reps)
            del A_s
            res["spmv_sigma4096"] = dict(ms=s_ms, kernel_ms=s_kms, gbs=B_spmv / s_ms / 1e6,
                                         kernel_gbs=B_spmv / s_kms / 1e6, max_rel_err_vs_fp64=s_err,
                                         stored_per_nnz=stored / nnz, gather_only_ms=s_gather)
            log(f"SELL-64 sigma 4096 SpMV: {s_ms * 1e3:.1f} us/SpMV, kernel {s_kms * 1e3:.1f} us "
                f"({B_spmv / s_kms / 1e6:.0f} GB/s), stored/nnz {stored / nnz:.3f}, rel err {s_err:.1e}")
        # the CPU-runnable configuration: plain CSR SpMV, n = 65,536, exactly 16
        # distinct uniform columns per row, fp32, 100 repetitions, each result
        # against the FP64 product (the matrix and x of the CPU leg)
        U = G.uniform_rows(65536, 16, seed=seed, dtype=np.float32)
        Au = U.to_csr(R0)
        xu_h = np.random.default_rng(2).random(U.n, dtype=np.float32)
        xu, yu = torch.from_numpy(xu_h).to(dev), torch.zeros(U.m, dtype=torch.float32, device=dev)
        ref_u = U.to_scipy().astype(np.float64) @ xu_h.astype(np.float64)
        for _ in range(3):
            gg.gpu_spmv_(yu, Au, xu)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(100):
            gg.gpu_spmv_(yu, Au, xu)
        e1.record()
        torch.cuda.synchronize()
        u_ms = e0.elapsed_time(e1) / 100
        u_err = float(np.max(np.abs(yu.cpu().numpy() - ref_u) / np.maximum(np.abs(ref_u), 1e-30)))
        assert u_err <= 1e-5, f"CSR n=65536 SpMV parity failed: max rel err {u_err}"
        B_u = U.nnz * 8 + 12 * U.m
        res["csr65536"] = dict(us_per_spmv=round(u_ms * 1e3, 2), x100_ms=round(u_ms * 100, 3),
                               gbs=round(B_u / u_ms / 1e6, 1), bytes=B_u, max_rel_err_vs_fp64=u_err)
        log(f"CSR n=65536 x16 SpMV: {u_ms * 1e3:.1f} us/SpMV ({B_u / u_ms / 1e6:.0f} GB/s "
            f"algorithmic; 8.5 MB, cache-resident), rel err {u_err:.1e}")
        del Au, xu, yu

    # ---------------------------------------------------------------- SpMM
    if args.only in ("", "spmm") and args.spmm_k > 0:
        kk = args.spmm_k
        Bm = torch.from_numpy(rng.random((n, kk), dtype=np.float32)).to(dev)   # row-major
        Cm = torch.zeros((n, kk), dtype=torch.float32, device=dev)
        gg.gpu_spmm_(Cm, A_f, Bm)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(3):
            gg.gpu_spmm_(Cm, A_f, Bm)
        e1.record()
        torch.cuda.synchronize()
        spmm_ms = e0.elapsed_time(e1) / 3
        # the reference's column-major (Julia) operand and result
        Bc, Cc = Bm.t().contiguous().t(), torch.zeros((kk, n), dtype=torch.float32, device=dev).t()
        gg.gpu_spmm_(Cc, A_f, Bc)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(3):
            gg.gpu_spmm_(Cc, A_f, Bc)
        e1.record()
        torch.cuda.synchronize()
        spmm_cm_ms = e0.elapsed_time(e1) / 3
        assert torch.allclose(Cc, Cm, rtol=1e-5, atol=1e-6), "column-major SpMM differs"
        del Bc, Cc
        B_spmm = nnz * 8 + (n + n) * kk * 4   # pattern and values once, B read once, C written once
        cols = np.random.default_rng(5).choice(kk, 3, replace=False)
        S64 = Rg.to_scipy().astype(np.float64)
        Bh = Bm.cpu().numpy().astype(np.float64)
        Ch = Cm.cpu().numpy()
        relerr = 0.0
        for j in cols:
            ref = S64 @ Bh[:, j]
            relerr = max(relerr, float(np.max(np.abs(Ch[:, j] - ref) / np.maximum(np.abs(ref), 1e-30))))
        assert relerr <= 1e-5, f"SpMM parity failed: max rel err {relerr}"
        del S64, Bh, Ch, Bm, Cm
        res["spmm"] = dict(k=kk, ms=spmm_ms, bytes=B_spmm, gbs=B_spmm / spmm_ms / 1e6,
                           colmajor_ms=spmm_cm_ms, max_rel_err_vs_fp64=relerr)
        log(f"SELL SpMM k={kk} (row-major B, C): {spmm_ms * 1e3:.1f} us ({B_spmm / spmm_ms / 1e6:.0f} GB/s "
            f"algorithmic), column-major {spmm_cm_ms * 1e3:.1f} us, rel err {relerr:.1e}")

    # ---------------------------------------------------------- batched BFS
    D = None
    if args.only in ("", "msbfs"):
        for _ in range(args.warmup):
            D = gg.multi_source_bfs(A_b, roots)
        torch.cuda.synchronize()
        lat = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            D = gg.multi_source_bfs(A_b, roots)
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t0)
        step_s = float(np.mean(lat))
        kms, kn = kernel_pass(_lib, ctx, lambda: gg.multi_source_bfs(A_b, roots), args.graph_steps)
        k = len(roots)
        B_ms = 4 * nnz + 2 * n * (k // 8)  # per level: pattern once + source masks in and out
        res["msbfs"] = dict(step_ms=step_s * 1e3, qps=k / step_s, p50_ms=pct(lat, 50) * 1e3,
                            p99_ms=pct(lat, 99) * 1e3, levels=kn, kernel_ms=kms / max(kn, 1),
                            bytes_per_launch=B_ms,
                            gather8_only_ms=gather_only_ms(torch, dev, A_b, 10, wide=True))
        log(f"one dense pass of the BFS layout's column streams with 8-byte gathers alone: "
            f"{res['msbfs']['gather8_only_ms'] * 1e3:.1f} us")
        log(f"batched BFS k={k}: {step_s * 1e3:.2f} ms/step, {k / step_s:.0f} q/s, "
            f"{kn} levels, kernel avg {kms / max(kn, 1) * 1e3:.1f} us")
        if args.wide_roots > 0:
            # the same traversal with more sources per launch (4 mask words per vertex)
            rw = np.random.default_rng(seed + args.wide_roots).choice(
                np.nonzero(deg)[0], args.wide_roots, replace=False).astype(np.int32) + 1
            Dw = gg.multi_source_bfs(A_b, rw)
            torch.cuda.synchronize()
            latw = []
            for _ in range(max(args.steps // 2, 2)):
                t0 = time.perf_counter()
                Dw = gg.multi_source_bfs(A_b, rw)
                torch.cuda.synchronize()
                latw.append(time.perf_counter() - t0)
            stepw = float(np.mean(latw))
            d1 = gg.bfs_distances(A_b, int(rw[-1]))
            assert torch.equal(d1, Dw[:, -1]), "wide batch differs from the plain BFS"
            del Dw
            res["msbfs_wide"] = dict(k=args.wide_roots, step_ms=stepw * 1e3,
                                     qps=args.wide_roots / stepw, p99_ms=pct(latw, 99) * 1e3)
            log(f"batched BFS k={args.wide_roots}: {stepw * 1e3:.2f} ms/step, "
                f"{args.wide_roots / stepw:.0f} q/s")

    # ------------------------------------------------------------ plain BFS
    if args.only in ("", "bfs"):
        pr = roots[:args.plain_roots]
        gg.bfs_distances(A_b, int(pr[0]))
        gg.bfs_distances_do(A_b, int(pr[0]), A_out=A_csr)
        torch.cuda.synchronize()
        lat_plain, lat_do = [], []
        mismatches = 0
        for i, s in enumerate(pr):
            t0 = time.perf_counter()
            d = gg.bfs_distances(A_b, int(s))
            torch.cuda.synchronize()
            lat_plain.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            d2 = gg.bfs_distances_do(A_b, int(s), A_out=A_csr)
            torch.cuda.synchronize()
            lat_do.append(time.perf_counter() - t0)
            if not torch.equal(d, d2):
                mismatches += 1
            if D is not None and not torch.equal(d, D[:, i]):
                mismatches += 1
        assert mismatches == 0, f"{mismatches} BFS results differ between plain, DO and batched"
        kms, kn = kernel_pass(_lib, ctx, lambda: gg.bfs_distances(A_b, int(pr[0])),
                              args.graph_steps)
        d0 = d if len(pr) == 1 else gg.bfs_distances(A_b, int(pr[0]))
        d0h = d0.cpu().numpy()
        reach = d0h != np.iinfo(np.int32).max
        comp_edges = int(deg[reach].sum()) // 2
        res["bfs"] = dict(plain_qps=len(pr) / sum(lat_plain), plain_p50_ms=pct(lat_plain, 50) * 1e3,
                          plain_p99_ms=pct(lat_plain, 99) * 1e3, do_qps=len(pr) / sum(lat_do),
                          do_p50_ms=pct(lat_do, 50) * 1e3, do_p99_ms=pct(lat_do, 99) * 1e3,
                          levels=kn, level_kernel_ms=kms / max(kn, 1),
                          gteps_plain=comp_edges / float(np.median(lat_plain)) / 1e9,
                          gteps_do=comp_edges / float(np.median(lat_do)) / 1e9)
        log(f"plain BFS: {len(pr) / sum(lat_plain):.0f} q/s p99 {pct(lat_plain, 99) * 1e3:.2f} ms; "
            f"DO BFS: {len(pr) / sum(lat_do):.0f} q/s p99 {pct(lat_do, 99) * 1e3:.2f} ms")

        res["bfs_root_levels"] = (int(pr[0]), d0h)

    # ------------------------------------------------------------ CPU baseline
    if args.only in ("", "cpu") and not args.no_cpu:
        res["cpu"] = cpu_leg(np, G, Rg, roots, seed, res.pop("bfs_root_levels", None))

    # ------------------------------------------------------- packed graphs
    if args.only in ("", "packed") and args.packed_graphs > 0:
        res["packed"] = packed_leg(args, np, torch, gg, G, R0, seed)

    # --------------------------------------------- autotuned, R-MAT scale 20
    if args.only in ("", "tuned") and args.tuned_scale > 0:
        res["tuned"] = tuned_leg(args, np, torch, gg, G, R0, ctx, _lib, seed)

    if serve_res is not None:
        res["serve"] = serve_res

    # ----------------------------------------------------------------- SSSP
    if args.only in ("", "sssp"):
        res["sssp"] = sssp_leg(args, np, torch, gg, G, R0, ctx, _lib, dev, Rg, roots, seed,
                               sigma, relabel)

    # ---------------------------------------------------------------- report
    ms = res.get("msbfs")
    bfs = res.get("bfs", {})
    spmv = res.get("spmv", {})
    if ms:
        value = ms["qps"]
        kernel_ms = ms["kernel_ms"]
        achieved = ms["bytes_per_launch"] / kernel_ms / 1e6
        roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=round(achieved / HBM_PEAK_GBS, 4),
                        traffic=measured_traffic(args, relabel, "msbfs_level"),
                        kernel="multi-source BFS level (msbfs_sell_kernel + direction switch + "
                               "top-down step + its share of the byte-level flush)",
                        bytes_per_launch=ms["bytes_per_launch"],
                        avg_launch_us=round(kernel_ms * 1e3, 2),
                        dense_pass_gather8_only_us=round(ms["gather8_only_ms"] * 1e3, 1))
    else:
        value, roofline = None, None
    spmv_roof = None
    if spmv:
        spmv_roof = dict(bound="hbm", achieved=round(spmv["kernel_gbs"], 1), peak=HBM_PEAK_GBS,
                         unit="GB/s", frac=round(spmv["kernel_gbs"] / HBM_PEAK_GBS, 4),
                         kernel="sell_spmv_kernel", bytes_per_launch=spmv["bytes"],
                         traffic=measured_traffic(args, relabel, "sell_spmv"),
                         avg_launch_us=round(spmv["kernel_ms"] * 1e3, 2),
                         spmv_call_gbs=round(spmv["gbs"], 1))
        if "gather_only_ms" in spmv:
            spmv_roof["gather_only_us"] = round(spmv["gather_only_ms"] * 1e3, 1)
        if ceiling:
            # against what the project's own streaming-read kernel reaches on this box
            spmv_roof["measured_read_gbs"] = ceiling["read_gbs"]
            spmv_roof["frac_of_measured_read"] = round(spmv["kernel_gbs"] / ceiling["read_gbs"], 4)
        if "spmv_sigma4096" in res:
            s4 = res["spmv_sigma4096"]
            spmv_roof["sigma4096"] = dict(achieved=round(s4["kernel_gbs"], 1),
                                          frac=round(s4["kernel_gbs"] / HBM_PEAK_GBS, 4),
                                          avg_launch_us=round(s4["kernel_ms"] * 1e3, 2),
                                          call_us=round(s4["ms"] * 1e3, 1),
                                          stored_per_nnz=round(s4["stored_per_nnz"], 4),
                                          gather_only_us=round(s4["gather_only_ms"] * 1e3, 1),
                                          max_rel_err_vs_fp64=s4["max_rel_err_vs_fp64"])
    if roofline and ceiling:
        roofline["measured_read_gbs"] = ceiling["read_gbs"]
        roofline["frac_of_measured_read"] = round(roofline["achieved"] / ceiling["read_gbs"], 4)
    cpu = res.get("cpu")
    cpu_base = None
    if cpu:
        B_spmv = nnz * 8 + 12 * n
        B_u = 65536 * 16 * 8 + 12 * 65536
        bf_iters = (res.get("sssp") or {}).get("rmat", {}).get("batched_iters")
        cpu_base = {
            "value": round(1.0 / cpu["bfs_s"], 3), "unit": "queries/s", "cores": cpu["threads"],
            "kind": "port",
            "sample": (f"one BFS (the pull loop of bfs.jl, levels) from the first root of R-MAT 22 on "
                       f"{cpu['threads']} threads (nproc {cpu['nproc']}); beside it one fp32 SpMV on "
                       "R-MAT 22, 100 CSR SpMVs of the n=65,536 uniform-rows matrix, one grid "
                       "Bellman-Ford source, and one iteration of the k=32 multi-source Bellman-Ford "
                       "scaled by the GPU run's iteration count"),
            "nproc": cpu["nproc"],
            "bfs_1thread_qps": round(1.0 / cpu["bfs_1t_s"], 3),
            "spmv_gbs": round(B_spmv / cpu["spmv_s"] / 1e9, 2),
            "spmv_1thread_gbs": round(B_spmv / cpu["spmv_1t_s"] / 1e9, 2),
            "csr65536_spmv_us": round(cpu["csr65536_x100_s"] * 1e4, 1),
            "csr65536_gbs": round(B_u * 100 / cpu["csr65536_x100_s"] / 1e9, 2),
            "csr65536_1thread_gbs": round(B_u * 100 / cpu["csr65536_x100_1t_s"] / 1e9, 2),
            "grid4096_bf_qps": round(1.0 / cpu["grid_bf_s"], 4),
            "grid4096_bf_iters": cpu["grid_bf_iters"],
            "rmat22_bf32_iter_s": round(cpu["rmat_bf32_iter_s"], 3),
            "rmat22_bf32_qps_est": (round(32 / (cpu["rmat_bf32_iter_s"] * bf_iters), 4)
                                    if bf_iters else None),
        }
    out = {
        "metric": "BFS/SSSP queries/s + p99 latency (batched vs plain); SpMV GB/s vs HBM peak",
        "value": None if value is None else round(value, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None if not ms else round(ms["step_ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bool (or-and over u64 source masks)",
        "data": "synthetic (seeded R-MAT generator)",
        "config": {"workload": f"rmat{args.scale} ef16 a.57 b.19 c.19 symmetrised permuted; "
                               f"batched BFS from {args.roots} random non-isolated roots "
                               "(one step = one k=64 or-and SpMM traversal); SELL-64 "
                               + ("degree-relabelled" if relabel else "sigma4096")
                               + (", row entries by column" if args.sort_entries else "")
                               + (", equal-degree rows in RCM order" if relabel and args.rcm
                                  else ""),
                   "n": n, "nnz": nnz, "roots": args.roots, "slice_size": 64, "sigma": sigma,
                   "relabel": relabel, "sorted_entries": args.sort_entries,
                   "rcm_ties": relabel and args.rcm, "heavy_threshold": args.heavy,
                   "graph_steps": args.graph_steps, "push_alpha": args.push_alpha,
                   "parallelism": "single GPU"},
        "p50_ms": None if not ms else round(ms["p50_ms"], 3),
        "p99_ms": None if not ms else round(ms["p99_ms"], 3),
        "plain_qps": round(bfs["plain_qps"], 1) if bfs else None,
        "plain_p50_ms": round(bfs["plain_p50_ms"], 3) if bfs else None,
        "plain_p99_ms": round(bfs["plain_p99_ms"], 3) if bfs else None,
        "do_qps": round(bfs["do_qps"], 1) if bfs else None,
        "do_p99_ms": round(bfs["do_p99_ms"], 3) if bfs else None,
        "batched_vs_plain": round(value / bfs["plain_qps"], 2) if (bfs and value) else None,
        "sell_build": res["sell_build"],
        "wide_batch": {k: (round(v, 3) if isinstance(v, float) else v)
                       for k, v in res["msbfs_wide"].items()} if "msbfs_wide" in res else None,
        "gteps_plain": round(bfs["gteps_plain"], 3) if bfs else None,
        "gteps_do": round(bfs["gteps_do"], 3) if bfs else None,
        "spmv_gbs": round(spmv["gbs"], 1) if spmv else None,
        "spmv_us": round(spmv["ms"] * 1e3, 1) if spmv else None,
        "spmv_roofline": spmv_roof,
        "csr65536": res.get("csr65536"),
        "spmm": {k: (round(v, 4) if isinstance(v, float) else v)
                 for k, v in res["spmm"].items()} if "spmm" in res else None,
        "roofline": roofline,
        "sssp": {name: {k: (round(v, 3) if isinstance(v, float) else v) for k, v in c.items()}
                 for name, c in res.get("sssp", {}).items()} or None,
        "packed": {k: (round(v, 3) if isinstance(v, float) else v)
                   for k, v in res["packed"].items()} if "packed" in res else None,
        "tuned": res.get("tuned"),
        "serve": res.get("serve"),
        "cpu_baseline": cpu_base,
        "hbm_ceiling": ceiling,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
