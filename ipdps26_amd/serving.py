"""Plain versus served execution of one request stream (count queries and update batches over many
hypergraphs), with per-request latency. Used by bench.py (--workload serving_c5) and the tests.

plain   one request at a time in arrival order, each a synchronous hm_count_static or
        hm_graph_update call on that graph's own handle;
served  the same requests submitted to an hm_server (slots run concurrently, one HIP stream per
        graph, requests on one graph keep their order, consecutive counts are coalesced);
        with maintain=True the server also keeps each graph's counts current from the exact
        update deltas, so count queries after a graph's first count need no launch; with
        batch_graphs=n a worker counts up to n ready graphs in one batched launch sequence.
hub_min_edges=n > 0 (both forms) counts and updates every graph of at least n hyperedges
hub-aware (hm_graph_set_hubs(g, -1)).
Both start from identical graphs and apply identical batches, so every result must match.
"""
from __future__ import annotations

import time

import numpy as np

from . import Graph, Server


def _latency_stats(lat_ms: list[float]) -> dict:
    a = np.asarray(lat_ms, dtype=np.float64)
    return {"p50_ms": float(np.percentile(a, 50)), "p99_ms": float(np.percentile(a, 99)),
            "mean_ms": float(a.mean())}


def run_plain(graphs, reqs, realtime: bool, hub_min_edges: int = 0) -> dict:
    handles = [Graph(off, nodes) for off, nodes in graphs]
    for h, (off, _) in zip(handles, graphs):
        if 0 < hub_min_edges <= off.size - 1:
            h.set_hubs(-1)
    for h in handles:  # allocate every graph's count scratch outside the measured stream
        h.count_static()
    results, lat = [], []
    t0 = time.perf_counter()
    for r in reqs:
        arrival = t0 + (r["t"] if realtime else 0.0)
        if realtime:
            delay = arrival - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
        g = handles[r["graph"]]
        if r["kind"] == "count":
            out = g.count_static().astype(np.int64)
        else:
            _, out = g.update(r["dels"], r["ins_off"], r["ins_nodes"])
        lat.append((time.perf_counter() - arrival) * 1e3)
        results.append(out)
    makespan = time.perf_counter() - t0
    for h in handles:
        h.close()
    return {"results": results, "makespan_s": makespan, "requests_per_s": len(reqs) / makespan,
            **_latency_stats(lat)}


def run_served(graphs, reqs, realtime: bool, slots: int = 4, budget_bytes: int = 64 << 30,
               maintain: bool = False, batch_graphs: int = 0, capture: bool = False,
               hub_min_edges: int = 0) -> dict:
    srv = Server(budget_bytes, slots, maintain_counts=maintain, batch_graphs=batch_graphs,
                 capture_updates=capture, hub_min_edges=hub_min_edges)
    gids = [srv.add_graph(off, nodes) for off, nodes in graphs]
    for t in [srv.submit_count(g) for g in gids]:  # same warm-up as the plain path
        srv.wait(t)
    tickets = []
    t0 = time.perf_counter()
    for r in reqs:
        if realtime:
            delay = t0 + r["t"] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
        if r["kind"] == "count":
            tickets.append(srv.submit_count(gids[r["graph"]]))
        else:
            tickets.append(srv.submit_update(gids[r["graph"]], r["dels"], r["ins_off"], r["ins_nodes"]))
    done = [srv.wait(t) for t in tickets]
    makespan = time.perf_counter() - t0
    stats = srv.stats()
    srv.close()
    return {"results": [d["counts"] for d in done], "makespan_s": makespan,
            "requests_per_s": len(reqs) / makespan, "coalesced": stats["coalesced"],
            "maintained": stats["maintained"], "batched_counts": stats["batched_counts"],
            "batch_launches": stats["batch_launches"],
            **_latency_stats([d["latency_ms"] for d in done])}


def count_triples(reqs, results) -> int:
    """Triples classified by the count queries of a stream."""
    return int(sum(int(r.sum()) for q, r in zip(reqs, results) if q["kind"] == "count"))
