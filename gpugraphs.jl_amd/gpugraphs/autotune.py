"""Autotuner for the layout and launch choices of one graph.

Times candidates on the device and keeps the fastest:
  - storage: CSR with 1..64 lanes per row, or SELL-C-sigma with slice size C,
    sigma windows or degree relabelling, heavy-row threshold and heavy
    segment length (gg_sell_build);
  - the direction-optimising BFS switch (alpha, beta) of bfs_distances_do;
  - the steps per captured graph launch of the level loops.
The workload decides what is timed: one plus-times SpMV, BFS levels from the
given roots, a multi-source BFS of up to 64 roots, Bellman-Ford from the first
root, or Bellman-Ford from up to 32 roots at once.  Results
are cached per graph signature (size, entries, degree profile) and workload,
in memory and optionally in a JSON file.
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .generators import HostCSR

WORKLOADS = ("spmv", "bfs", "msbfs", "sssp", "msssp")


@dataclass
class TuneResult:
    workload: str
    best: dict
    best_ms: float
    table: list = field(default_factory=list)   # (candidate, ms) for every candidate timed


def default_candidates(H: HostCSR):
    deg = H.degrees()
    mx = int(deg.max()) if len(deg) else 0
    c = [dict(layout="csr", lanes=L) for L in (1, 4, 16, 64)]
    for heavy in (256, 1024, 4096):
        if heavy < mx:
            for seg in (4096, 65536):
                c.append(dict(layout="sell", slice_size=64, sigma=1, relabel=True, heavy=heavy,
                              seg=seg))
    c.append(dict(layout="sell", slice_size=64, sigma=1, relabel=True, heavy=0, seg=4096))
    if 1024 < mx:  # each row's entries by increasing internal column (light hubs first, heavy rows last)
        c.append(dict(layout="sell", slice_size=64, sigma=1, relabel=True, heavy=1024, seg=4096,
                      sort=True))
    c.append(dict(layout="sell", slice_size=64, sigma=4096, relabel=False,
                  heavy=256 if 256 < mx else 0, seg=4096))
    c.append(dict(layout="sell", slice_size=32, sigma=1, relabel=True,
                  heavy=1024 if 1024 < mx else 0, seg=4096))
    if H.m != H.n:   # relabelling needs a square matrix
        c = [x for x in c if not x.get("relabel")]
    return c


def signature(H: HostCSR, workload: str) -> str:
    deg = H.degrees()
    hist = np.histogram(np.log2(deg + 1), bins=8, range=(0, 24))[0] if len(deg) else []
    return f"{workload}:{H.m}x{H.n}:{H.nnz}:{','.join(str(int(x)) for x in hist)}"


def build(H: HostCSR, backend, cand: dict):
    if cand["layout"] == "csr":
        return H.to_csr(backend)
    return H.to_sell(backend, cand["slice_size"], cand["sigma"], cand["heavy"], cand["seg"],
                     cand["relabel"], cand.get("sort", False))


def _runner(workload, A, H, roots, torch, gg):
    from .algorithms import multi_source_bfs, shortest_path_
    dev = torch.device("cuda", A.backend.device)
    if workload == "spmv":
        x = torch.rand(H.n, dtype=torch.float32, device=dev)
        y = torch.zeros(H.m, dtype=torch.float32, device=dev)
        return lambda: gg.gpu_spmv_(y, A, x)
    if workload == "bfs":
        return lambda: [gg.bfs_distances(A, int(r)) for r in roots]
    if workload == "msbfs":
        r = np.asarray(roots[:64], np.int32)
        return lambda: multi_source_bfs(A, r)
    dt = torch.float32 if H.nzval.dtype == np.float32 else torch.int32
    if workload == "msssp":
        r = np.asarray(roots[:32], np.int32)
        D = torch.empty(len(r) * H.m, dtype=dt, device=dev).view(len(r), H.m).T
        return lambda: shortest_path_(A, r, D)
    d = torch.empty(H.m, dtype=dt, device=dev)
    return lambda: shortest_path_(A, int(roots[0]), d)


class Autotuner:
    def __init__(self, backend, cache_path: str | None = None, reps: int = 3):
        self.backend = backend
        self.reps = reps
        self.cache_path = cache_path
        self.cache = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.cache = json.load(f)

    def _save(self):
        if self.cache_path:
            with open(self.cache_path, "w") as f:
                json.dump(self.cache, f, indent=1)

    def _time(self, fn, torch):
        fn()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(self.reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best * 1e3

    def tune_layout(self, H: HostCSR, workload: str = "spmv", roots=None, candidates=None):
        """Fastest storage for `workload` on H (the A_T of the graph; SpMV
        uses H's values as fp32, Bellman-Ford H's values)."""
        import torch
        import gpugraphs as gg
        if workload not in WORKLOADS:
            raise _lib.ArgumentError(f"workload must be one of {WORKLOADS}")
        if workload != "spmv" and roots is None:
            raise _lib.ArgumentError("graph workloads need roots")
        key = signature(H, workload)
        if key in self.cache and candidates is None:
            c = self.cache[key]
            return TuneResult(workload, c["best"], c["best_ms"], c.get("table", []))
        Hs = H
        if workload == "spmv" and H.nzval.dtype != np.float32:
            Hs = HostCSR(H.m, H.n, H.rowptr, H.colval, np.ones(H.nnz, np.float32))
        ctx = _lib.context(self.backend.device)
        table = []
        for cand in candidates or default_candidates(H):
            A = build(Hs, self.backend, cand)
            _lib.set_csr_lanes(ctx, cand.get("lanes", 0))
            try:
                ms = self._time(_runner(workload, A, Hs, roots, torch, gg), torch)
            finally:
                _lib.set_csr_lanes(ctx, 0)
            table.append((cand, ms))
            del A
        best, best_ms = min(table, key=lambda t: t[1])
        self.cache[key] = dict(best=best, best_ms=best_ms, table=table)
        self._save()
        return TuneResult(workload, best, best_ms, table)

    def tune_direction(self, A_T, A_out, roots, alphas=(4.0, 15.0, 64.0), betas=(6.0, 18.0, 64.0)):
        """(alpha, beta) of the direction-optimising BFS with the least total
        time over the roots."""
        import torch
        import gpugraphs as gg
        table = []
        for a in alphas:
            for b in betas:
                ms = self._time(lambda: [gg.bfs_distances_do(A_T, int(r), A_out=A_out, alpha=a,
                                                             beta=b) for r in roots], torch)
                table.append(((a, b), ms))
        (a, b), ms = min(table, key=lambda t: t[1])
        return TuneResult("bfs_do", dict(alpha=a, beta=b), ms, table)

    def tune_push_threshold(self, fn, alphas=(0.0, 4.0, 16.0, 64.0)):
        """Direction switch of the multi-source BFS (gg_ctx_set_push_threshold)
        with the least time for fn(); 0 keeps every level bottom-up."""
        import torch
        ctx = _lib.context(self.backend.device)
        table = []
        for a in alphas:
            _lib.set_push_threshold(ctx, a)
            table.append((a, self._time(fn, torch)))
        a, ms = min(table, key=lambda t: t[1])
        _lib.set_push_threshold(ctx, a)
        return TuneResult("push_threshold", dict(alpha=a), ms, table)

    def tune_graph_steps(self, fn, steps=(0, 2, 8, 32)):
        """Steps per captured graph launch with the least time for fn()."""
        import torch
        ctx = _lib.context(self.backend.device)
        table = []
        for s in steps:
            _lib.set_graph_steps(ctx, s)
            table.append((s, self._time(fn, torch)))
        s, ms = min(table, key=lambda t: t[1])
        _lib.set_graph_steps(ctx, s)
        return TuneResult("graph_steps", dict(steps=s), ms, table)
