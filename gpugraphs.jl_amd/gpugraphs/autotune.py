"""Autotuner for the layout and launch choices of one graph.

Times candidates on the device and keeps the fastest:
  - storage: CSR with 1..64 lanes per row, or SELL-C-sigma with slice size C,
    sigma windows or degree relabelling, heavy-row threshold and heavy
    segment length (gg_sell_build);
  - the direction-optimising BFS switch (alpha, beta) of bfs_distances_do;
  - the top-down switch of the multi-source BFS and the sparse-iteration
    switch of the k-source Bellman-Ford;
  - the steps per captured graph launch of the level loops.
The workload decides what is timed: one plus-times SpMV, BFS levels from the
given roots, a multi-source BFS of up to 64 roots, Bellman-Ford from the first
root, or Bellman-Ford from up to 32 roots at once.  The layout and push/pull
searches run in the library (gg_tune_layout, gg_tune_push_threshold,
gg_tune_sssp_threshold: every
candidate is built on the device from the graph's CSR and timed with HIP
events), so a caller of the C ABI tunes the same way.  Results are cached per
graph signature (size, entries, degree profile) and workload, in memory and
optionally in a JSON file.
"""
from __future__ import annotations

import ctypes as C
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


_WORKLOAD_CODE = {"spmv": 1, "bfs": 2, "msbfs": 3, "sssp": 4, "msssp": 5}


def _to_native(cand: dict) -> _lib.TuneCandidate:
    c = _lib.TuneCandidate()
    if cand["layout"] == "csr":
        c.format, c.csr_lanes = _lib.CSR, int(cand.get("lanes", 0))
    else:
        c.format = _lib.SELL
        c.slice_size, c.sigma = int(cand["slice_size"]), int(cand["sigma"])
        c.heavy_threshold, c.heavy_seg_len = int(cand["heavy"]), int(cand["seg"])
        c.sell_flags = ((_lib.SELL_RELABEL if cand.get("relabel") else 0)
                        | (_lib.SELL_SORT_ENTRIES if cand.get("sort") else 0))
    return c


def _from_native(c: _lib.TuneCandidate) -> dict:
    if c.format == _lib.CSR:
        return dict(layout="csr", lanes=int(c.csr_lanes))
    d = dict(layout="sell", slice_size=int(c.slice_size), sigma=int(c.sigma),
             relabel=bool(c.sell_flags & _lib.SELL_RELABEL), heavy=int(c.heavy_threshold),
             seg=int(c.heavy_seg_len))
    if c.sell_flags & _lib.SELL_SORT_ENTRIES:
        d["sort"] = True
    return d


def default_candidates(H: HostCSR):
    """The native default candidate set (gg_tune_default_candidates) as dicts."""
    deg = H.degrees()
    d = _lib.GGMatrix()
    d.format, d.m, d.n, d.nnz = _lib.CSR, H.m, H.n, H.nnz
    buf = (_lib.TuneCandidate * 64)()
    n = C.c_int64()
    _lib.check(_lib.lib().gg_tune_default_candidates(C.byref(d), int(deg.max()) if len(deg) else 0,
                                                     buf, 64, C.byref(n)), "default_candidates")
    return [_from_native(buf[i]) for i in range(n.value)]


def signature(H: HostCSR, workload: str) -> str:
    deg = H.degrees()
    hist = np.histogram(np.log2(deg + 1), bins=8, range=(0, 24))[0] if len(deg) else []
    return f"{workload}:{H.m}x{H.n}:{H.nnz}:{','.join(str(int(x)) for x in hist)}"


def build(H: HostCSR, backend, cand: dict):
    if cand["layout"] == "csr":
        return H.to_csr(backend)
    return H.to_sell(backend, cand["slice_size"], cand["sigma"], cand["heavy"], cand["seg"],
                     cand["relabel"], cand.get("sort", False))


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
        # the native tuner builds every candidate on the device from the CSR
        # and times the workload (gg_tune_layout); SpMV times fp32 ones when
        # the values are of another type
        cands = candidates or default_candidates(H)
        arr = (_lib.TuneCandidate * len(cands))(*[_to_native(c) for c in cands])
        A = H.to_csr(self.backend)
        d = A.descriptor()
        r = np.ascontiguousarray(roots if roots is not None else [1], np.int32)
        best = _lib.TuneCandidate()
        ctx = _lib.context(self.backend.device)
        _lib.check(_lib.lib().gg_tune_layout(ctx, C.byref(d), _WORKLOAD_CODE[workload],
                                             r.ctypes.data, len(r), arr, len(cands), self.reps,
                                             C.byref(best)), "tune_layout")
        table = [(cands[i], float(arr[i].ms)) for i in range(len(cands)) if arr[i].ms >= 0]
        best_d, best_ms = min(table, key=lambda t: t[1])
        self.cache[key] = dict(best=best_d, best_ms=best_ms, table=table)
        self._save()
        return TuneResult(workload, best_d, best_ms, table)

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

    def tune_push_threshold(self, fn, alphas=(0.0, 4.0, 16.0, 64.0), roots=None):
        """Direction switch of the multi-source BFS (gg_ctx_set_push_threshold)
        with the least time; 0 keeps every level bottom-up.  fn is a matrix
        (A_T, timed natively with gg_tune_push_threshold on up to 64 roots)
        or a callable timed as it is."""
        import torch
        ctx = _lib.context(self.backend.device)
        if not callable(fn):
            r = np.ascontiguousarray(roots[:64], np.int32)
            al = np.ascontiguousarray(alphas, np.float64)
            ms = np.zeros(len(al), np.float64)
            best = C.c_double()
            d = fn.descriptor()
            _lib.check(_lib.lib().gg_tune_push_threshold(ctx, C.byref(d), r.ctypes.data, len(r),
                                                         al.ctypes.data, ms.ctypes.data, len(al),
                                                         self.reps, C.byref(best)),
                       "tune_push_threshold")
            return TuneResult("push_threshold", dict(alpha=best.value), float(ms.min()),
                              list(zip(al.tolist(), ms.tolist())))
        table = []
        for a in alphas:
            _lib.set_push_threshold(ctx, a)
            table.append((a, self._time(fn, torch)))
        a, ms = min(table, key=lambda t: t[1])
        _lib.set_push_threshold(ctx, a)
        return TuneResult("push_threshold", dict(alpha=a), ms, table)

    def tune_sssp_threshold(self, A_T, roots, alphas=(0.0, 16.0, 64.0, 256.0, 1024.0)):
        """Sparse-iteration switch of the k-source Bellman-Ford
        (gg_ctx_set_sssp_threshold) with the least time for
        shortest_path(A_T, roots[:32]), timed natively with
        gg_tune_sssp_threshold; the context keeps the best."""
        ctx = _lib.context(self.backend.device)
        r = np.ascontiguousarray(roots[:32], np.int32)
        al = np.ascontiguousarray(alphas, np.float64)
        ms = np.zeros(len(al), np.float64)
        best = C.c_double()
        d = A_T.descriptor()
        _lib.check(_lib.lib().gg_tune_sssp_threshold(ctx, C.byref(d), r.ctypes.data, len(r),
                                                     al.ctypes.data, ms.ctypes.data, len(al),
                                                     self.reps, C.byref(best)),
                   "tune_sssp_threshold")
        return TuneResult("sssp_threshold", dict(alpha=best.value), float(ms.min()),
                          list(zip(al.tolist(), ms.tolist())))

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
