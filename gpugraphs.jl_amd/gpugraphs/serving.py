"""Continuous-batching BFS / SSSP server over HBM-resident graphs.

A thin wrapper of the native server (gg_server_* in include/gpugraphs.h):
graphs are registered once, queries are submitted from any thread, the
native scheduler batches them into the bitmask BFS and min-plus engines of
their graph, and results come back per query, identical to ``bfs_distances``
/ ``shortest_path`` from the same source.
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import EngineCounters, QueryInfo, ServerCounters, check, lib
from .algorithms import dist_dtype

BFS, SSSP = 1, 2


@dataclass
class Completion:
    request_id: int
    graph: int
    kind: int
    source: int
    status: int
    arrival_s: float
    admit_s: float
    complete_s: float

    @property
    def latency_s(self):
        return self.complete_s - self.arrival_s


class GraphServer:
    """Query server over registered graphs (A_T matrices on one device)."""

    def __init__(self, budget_bytes: int, device: int = 0):
        import torch
        torch.cuda.init()
        self._L = lib()
        h = C.c_void_p()
        check(self._L.gg_server_create(device, int(budget_bytes), C.byref(h)), "gg_server_create")
        self._h = h
        self._graphs = []       # keep matrices (and their descriptors) alive
        self._meta = []         # (n, result dtype per kind)
        self._started = False
        _lib.register_server(self)

    def add_graph(self, A_T) -> int:
        d = A_T.descriptor()
        gid = C.c_int()
        check(self._L.gg_server_add_graph(self._h, C.byref(d), C.byref(gid)), "add_graph")
        self._graphs.append((A_T, d))
        self._meta.append((A_T.m, {BFS: np.dtype(np.int32), SSSP: np.dtype(dist_dtype(A_T))}))
        return gid.value

    def set_pacing(self, factor: float):
        """Step pacing factor (gg_server_set_pacing; 0 steps whenever a slot
        is busy)."""
        check(self._L.gg_server_set_pacing(self._h, float(factor)), "set_pacing")

    def start(self):
        check(self._L.gg_server_start(self._h), "start")
        self._started = True

    def submit(self, graph: int, kind: int, source: int) -> int:
        rid = C.c_int64()
        check(self._L.gg_server_submit(self._h, graph, kind, int(source), C.byref(rid)), "submit")
        return rid.value

    def poll(self, max_items: int = 4096, wait_s: float = 0.0):
        buf = (QueryInfo * max_items)()
        n = C.c_int64()
        check(self._L.gg_server_poll(self._h, buf, max_items, C.byref(n), wait_s), "poll")
        return [Completion(b.request_id, b.graph, b.kind, b.source, b.status, b.arrival_s,
                           b.admit_s, b.complete_s) for b in buf[:n.value]]

    def fetch(self, request_id: int, graph: int, kind: int, wait_s: float = 60.0) -> np.ndarray:
        n, dts = self._meta[graph]
        out = np.empty(n, dts[kind])
        check(self._L.gg_server_fetch(self._h, request_id, out.ctypes.data, out.nbytes, wait_s),
              "fetch")
        return out

    def release(self, request_id: int, wait_s: float = 60.0):
        check(self._L.gg_server_release(self._h, request_id, wait_s), "release")

    def stats(self) -> dict:
        c = ServerCounters()
        check(self._L.gg_server_stats(self._h, C.byref(c)), "stats")
        return {f: getattr(c, f) for f, _ in ServerCounters._fields_}

    def engine_stats(self, graph: int, kind: int) -> dict:
        """Counters of one engine of a graph (gg_server_engine_stats): steps,
        busy_slot_steps, finished, slots, device_s."""
        c = EngineCounters()
        check(self._L.gg_server_engine_stats(self._h, graph, kind, C.byref(c)), "engine_stats")
        return {f: getattr(c, f) for f, _ in EngineCounters._fields_ if f != "reserved"}

    def stop(self):
        check(self._L.gg_server_stop(self._h), "stop")

    def close(self):
        if self._h:
            self._L.gg_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self, _finalizing=_lib.finalizing):
        # at interpreter shutdown the atexit teardown has closed every live
        # server; no native call is made once finalisation has begun.  The
        # check is bound at definition time: module globals may already be
        # None when a late collection runs this.
        if not getattr(self, "_h", None) or _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def run_poisson(server: GraphServer, graphs, rate_qps: float, duration_s: float, seed: int = 0,
                sssp_fraction: float = 0.5, keep_every: int = 0):
    """Submit Poisson arrivals (rate_qps) for duration_s: each query is BFS or
    SSSP (sssp_fraction) from a uniform random source on a uniform random
    graph of ``graphs``, a list of (graph_id, n, sources_pool or None) or of
    (bfs_graph_id, sssp_graph_id, n, sources_pool or None) when BFS and SSSP
    queries of one graph go to two registrations (pattern and weights).  A
    collector thread polls completions and fetches or releases each result;
    every keep_every-th result is kept.  Returns (completions, wall_s,
    results by request id, submitted (request id, graph id, kind, source)).

    The submitting thread sleeps until each arrival time, so the collector is
    not kept waiting for the interpreter lock: a busy-wait here held the lock
    for most of the run, and results the collector had not yet released
    filled the server's result arena, which then held back admissions."""
    rng = np.random.default_rng(seed)
    done, results = [], {}
    submitted = []
    lock = threading.Lock()
    finished = threading.Event()

    peaks = dict(arena_used=0, waiting=0)
    server.last_run_peaks = peaks
    t_sample = [0.0]

    def collector():
        while True:
            batch = server.poll(wait_s=0.01)
            now = time.perf_counter()
            if now - t_sample[0] > 0.02:  # admission state, sampled every 20 ms
                t_sample[0] = now
                st = server.stats()
                peaks["arena_used"] = max(peaks["arena_used"], st["arena_used"])
                peaks["waiting"] = max(peaks["waiting"], st["waiting"])
            for c in batch:
                if c.status == 0:
                    if keep_every and c.request_id % keep_every == 0:
                        results[c.request_id] = server.fetch(c.request_id, c.graph, c.kind)
                    else:
                        server.release(c.request_id)
                with lock:
                    done.append(c)
            with lock:
                if finished.is_set() and len(done) >= len(submitted):
                    return

    th = threading.Thread(target=collector, daemon=True)
    th.start()
    t0 = time.perf_counter()
    t_next = t0
    while True:
        t_next += rng.exponential(1.0 / rate_qps)
        if t_next - t0 > duration_s:
            break
        wait = t_next - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        spec = graphs[rng.integers(len(graphs))]
        if len(spec) == 4:
            gb, gs, n, pool = spec
        else:
            gb, n, pool = spec
            gs = gb
        src = int(pool[rng.integers(len(pool))]) if pool is not None else int(rng.integers(n)) + 1
        kind = SSSP if rng.random() < sssp_fraction else BFS
        gid = gs if kind == SSSP else gb
        rid = server.submit(gid, kind, src)
        with lock:
            submitted.append((rid, gid, kind, src))
    finished.set()
    th.join()
    wall = time.perf_counter() - t0
    return done, wall, results, submitted
