#!/usr/bin/env python3
"""Throughput of exact h-motif counting on one MI355X.

Workloads (inputs are built and copied to HBM before the timed region):
  temporal_c4 (default)  THyMe's 96 temporal h-motifs on the temporal workload: 200,000 temporal
                         hyperedges over 100,000 nodes (sizes ~k^-2.2 on [2,20], Zipf s=1.2,
                         seed 4), int64 microsecond timestamps (Exp(1 s) gaps, 1,200 bursts of 50
                         with Exp(10 ms) gaps), window delta = 3,600 s. A step is one
                         hm_count_temporal call (window bounds, temporal projection, counting).
  static_c1              MoCHy's 30 static bins on 5,000 hyperedges over 2,000 nodes (sizes
                         ~k^-2.5 on [2,25], Zipf s=1.2, seed 1). A step is one hm_count_static call.
  dynamic_c3             the dynamic workload's generator and batch recipe at a reduced base size
                         (20,000 hyperedges over 8,000 nodes: c3's V/E and batch/E ratios). A step
                         is one batch: hm_graph_update with its exact delta (store update + delta).
                         value = triples touched by the batches (removed + added) per second.
  serving_c5             plain versus served request streams over 64 reduced graphs (p50/p99).
  static_c3_base         MoCHy's 30 static bins of the dynamic workload's full 500,000-hyperedge
                         base graph (3.16e15 triples, its largest node in 198,573 hyperedges) by
                         the hub-aware count (hm_graph_set_hubs, hub count from the cost model); a
                         step takes ~18 s, so this workload defaults to one step, no warmup.
The full static 1M-node / 2M-hyperedge graph counts exactly with the hub-aware count in ~21 min of
GPU time, split into time-bounded parts (hm_graph_count_hub_part; tests/golden/make_static_c2.py,
tests/golden/static_c2_hub.json): beyond a bench step. The temporal workload is the largest
full-size configuration whose step is short enough for repeated timing; the default line adds the
static counts of c1 and of a 60,000-hyperedge graph of the 2M-hyperedge graph's family by the
hub-aware count (config.static_hub).

value = h-motif instances classified per second = (sum of the bins) / step time.
roofline = the dominant kernel (k_tcount, k_tri or k_delta), bound "hbm": achieved = the HBM bytes
one launch moves, read by rocprofv3 (FETCH_SIZE x 2 + WRITE_SIZE x 1, the factors a plain 1 GiB
device copy needs to read its own size; profiles/traffic_<workload>.txt, same build), divided by
the kernel's average duration from HIP events on its stream in this run; frac = achieved / 8 TB/s.
The cost model of the counting path (8 B per directed projection pair + 8 B per neighbour-row
entry read for closed triples + 8 B per bin) is reported beside it as model_bytes: most of those
bytes come from LDS, L2 and the Infinity Cache, so it is not an HBM lower bound. valu = the
kernel's SQ_INSTS_VALU per launch (profiles/counters_<kernel>_<workload>.md) per second of its
duration here, against the issue rate of a VALU-only calibration kernel measured under the same
counter (profiles/valu_probe.txt; bench.py --valu-probe).
cpu_baseline = the reference program timed on this host on a bounded prefix of the same input:
THyMe+ main_thymeP (temporal) or MoCHy main_exact_par (static), built from the reference sources
into oracle/_ref, with `nproc` threads (cores and SMT of the host stated beside it).
serving (default run) = a shortened plain-versus-served request stream over the reduced serving
set: requests/s, p50/p99, results identical request by request.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import ipdps26_amd as hm  # noqa: E402
from ipdps26_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
DESC = {
    "temporal_c4": "temporal_c4: 200,000 temporal hyperedges over 100,000 nodes, sizes ~k^-2.2 on "
                   "[2,20], Zipf s=1.2, seed 4, int64 us timestamps, delta=3600 s",
    "static_c1": "static_c1: 5,000 hyperedges over 2,000 nodes, sizes ~k^-2.5 on [2,25], "
                 "Zipf s=1.2, seed 1",
    "dynamic_c3": "dynamic_c3 (reduced): base graph of the dynamic generator family (sizes ~k^-2.2 "
                  "on [2,30], Zipf s=1.1, seed 3) with 20,000 hyperedges over 8,000 nodes; batch b: "
                  "280 fresh generator hyperedges (seed 3000+b) and 120 deletes of uniformly chosen "
                  "live hyperedges (c3's 7,000 / 3,000 per 500,000 scaled to E); steps = batches",
    "serving_c5": "serving_c5 (reduced): 64 graphs of the static_c2 family, E log-uniform on "
                  "[1k, 6k], V = E/2, seed 5; 400 requests, half count queries, half update "
                  "batches of 0.5-5% of E (70/30 insert/delete), seed 5000",
    "static_c3_base": "static_c3_base: the dynamic workload's base graph at full size, 500,000 "
                      "hyperedges over 200,000 nodes, sizes ~k^-2.2 on [2,30], Zipf s=1.1, seed 3; "
                      "hub-aware count (hubs from the cost model)",
}
WHY_DEFAULT = ("the largest full-size configuration with a step short enough for repeated timing: "
               "the 2M-hyperedge static graph (one node in 780,282 hyperedges, 1.96e17 triples) "
               "counts exactly with the hub-aware count in ~21 min of GPU time, split into two "
               "time-bounded parts (tests/golden/static_c2_hub.json; 22 and 23 hubs agree bin for "
               "bin), beyond a bench step; the 500k-hyperedge dynamic base counts exactly in ~18 s "
               "per step (--workload static_c3_base)")
# Bounded CPU samples: a prefix of the same input sized for ~10-30 s of reference-program time.
CPU_SAMPLE_EDGES = {"temporal_c4": 700, "static_c1": 4000, "dynamic_c3": 4600}
DEFAULT_STEPS = {"temporal_c4": 3, "static_c1": 20, "dynamic_c3": 10, "serving_c5": 1, "static_c3_base": 1}
DYN_BASE_EDGES, DYN_BASE_NODES = 20_000, 8_000


def cpu_threads() -> int:
    """Threads for the reference CPU programs: `nproc` (the CPUs this process may run on, as
    limited by its affinity mask and OMP_NUM_THREADS)."""
    try:
        return max(1, int(subprocess.run(["nproc"], capture_output=True, text=True, check=True).stdout))
    except (OSError, ValueError, subprocess.CalledProcessError):
        return len(os.sched_getaffinity(0)) or 1


def host_cpu() -> dict:
    """Logical CPUs, physical cores and SMT of the CPUs this process may run on (sysfs)."""
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology"
        try:
            pkg = open(f"{base}/physical_package_id").read().strip()
            core = open(f"{base}/core_id").read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add((str(c), str(c)))
    model = ""
    try:
        model = next((l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo")
                      if l.startswith("model name")), "")
    except OSError:
        pass
    return {"logical_cpus_allowed": len(cpus), "physical_cores_allowed": len(cores),
            "smt_threads_per_core": round(len(cpus) / max(1, len(cores)), 2),
            "online_logical_cpus": os.cpu_count(), "model": model}


def cpu_baseline_static(off, nodes) -> dict:
    k = CPU_SAMPLE_EDGES["static_c1"]
    soff, snodes = off[: k + 1], nodes[: off[k]]
    with hm.Graph(soff, snodes) as g:  # MoCHy prints 26 of 30 bins; the total comes from here
        triples = int(g.count_static().astype(np.int64).sum())
    threads = cpu_threads()
    ref = os.path.join(ROOT, "oracle", "_ref", "mochy_exact_par")
    with tempfile.TemporaryDirectory() as td:
        csv = os.path.join(td, "sample.csv")
        W.write_mochy_csv(csv, soff, snodes)
        if os.path.exists(ref):
            t0 = time.perf_counter()
            out = subprocess.run([ref, str(threads), csv], capture_output=True, text=True,
                                 check=True, timeout=900).stdout
            wall = time.perf_counter() - t0
            secs = float(re.search(r"TOTAL_RUNTIME_SECONDS ([0-9.eE+-]+)", out).group(1))
            kind, prog = "reference", "MoCHy main_exact_par"
        else:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import oracle_lib

            t0 = time.perf_counter()
            oracle_lib.static_mochy(soff, snodes, threads)
            secs = wall = time.perf_counter() - t0
            kind, prog = "port", "oracle restatement of MoCHy-E"
    return {"value": triples / secs, "unit": "triples/s", "cores": threads, "kind": kind,
            "sample": f"first {k} of {off.size - 1} hyperedges of static_c1 ({triples} triples); "
                      f"{prog} TOTAL_RUNTIME_SECONDS={secs:.2f} (wall {wall:.2f} s)"}


def cpu_baseline_recount(off, nodes) -> dict:
    """The reference's per-batch CPU path for dynamic graphs: a full MoCHy recount (timed on a
    bounded prefix of the updated graph, as triples recounted per second)."""
    k = CPU_SAMPLE_EDGES["dynamic_c3"]
    sizes = np.diff(off)
    keep = np.nonzero(sizes > 0)[0][:k]  # deleted rows are empty; MoCHy's reader takes none
    soff = np.concatenate([[0], np.cumsum(sizes[keep])]).astype(np.int64)
    snodes = np.concatenate([nodes[off[e]:off[e + 1]] for e in keep]).astype(np.int32)
    with hm.Graph(soff, snodes) as g:
        triples = int(g.count_static().astype(np.int64).sum())
    threads = cpu_threads()
    ref = os.path.join(ROOT, "oracle", "_ref", "mochy_exact_par")
    with tempfile.TemporaryDirectory() as td:
        csv = os.path.join(td, "sample.csv")
        W.write_mochy_csv(csv, soff, snodes)
        if os.path.exists(ref):
            out = subprocess.run([ref, str(threads), csv], capture_output=True, text=True,
                                 check=True, timeout=900).stdout
            secs = float(re.search(r"TOTAL_RUNTIME_SECONDS ([0-9.eE+-]+)", out).group(1))
            kind, prog = "reference", "MoCHy main_exact_par"
        else:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import oracle_lib

            t0 = time.perf_counter()
            oracle_lib.static_mochy(soff, snodes, threads)
            secs = time.perf_counter() - t0
            kind, prog = "port", "oracle restatement of MoCHy-E"
    return {"value": triples / secs, "unit": "triples/s", "cores": threads, "kind": kind,
            "sample": f"full recount (the reference's per-batch path) of the first {k} live of "
                      f"{int((np.diff(off) > 0).sum())} hyperedges of the updated dynamic_c3 graph ({triples} "
                      f"triples); {prog} TOTAL_RUNTIME_SECONDS={secs:.2f}"}


def cpu_baseline_temporal(off, nodes, times, delta) -> dict:
    k = CPU_SAMPLE_EDGES["temporal_c4"]
    soff, snodes, stimes = off[: k + 1], nodes[: off[k]], times[:k]
    threads = cpu_threads()
    ref = os.path.join(ROOT, "oracle", "_ref", "thymeP")
    with tempfile.TemporaryDirectory() as td:
        if os.path.exists(ref):
            tsv, res = os.path.join(td, "sample.tsv"), os.path.join(td, "res.txt")
            W.write_thyme_tsv(tsv, soff, snodes, stimes)
            t0 = time.perf_counter()
            subprocess.run([ref, "x", str(delta), str(threads), tsv, res], capture_output=True,
                           check=True, timeout=900)
            secs = time.perf_counter() - t0
            inst = sum(int(l.split("\t")[1]) for l in open(res).read().split("\n")[1:97])
            kind, prog = "reference", "THyMe+ main_thymeP"
        else:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import oracle_lib

            t0 = time.perf_counter()
            inst = int(oracle_lib.thyme_temporal(soff, snodes, stimes, delta).sum())
            secs = time.perf_counter() - t0
            kind, prog = "port", "oracle restatement of THyMe+"
    return {"value": inst / secs, "unit": "instances/s", "cores": threads, "kind": kind,
            "sample": f"first {k} of {off.size - 1} temporal hyperedges of temporal_c4, "
                      f"delta=3600 s ({inst} instances); {prog} wall time {secs:.2f} s"}


def _profile_kv(name: str) -> dict:
    """key: value lines of a committed rocprofv3 summary under profiles/ ({} when absent)."""
    path = os.path.join(ROOT, "profiles", name)
    out = {}
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                key, sep, val = line.partition(":")
                if sep:
                    out[key.strip()] = val.strip()
    return out


def measured_traffic(workload: str):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 summary."""
    v = _profile_kv(f"traffic_{workload}.txt").get("hbm_bytes_per_launch")
    return int(v) if v else None


def measured_valu(kernel: str, workload: str):
    """SQ_INSTS_VALU of one launch of `kernel` (profiles/counters_<kernel>_<workload>.md)."""
    path = os.path.join(ROOT, "profiles", f"counters_{kernel}_{workload}.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"^\|\s*SQ_INSTS_VALU\s*\|\s*([0-9.eE+]+)\s*\|", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def roofline(workload: str, r: dict) -> dict:
    kms = r["kernel_ms"]
    traffic = measured_traffic(workload)
    achieved = traffic / (kms * 1e-3) / 1e9 if traffic else None
    out = {"bound": "hbm", "kernel": r["kernel"], "achieved": achieved, "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
           "kernel_ms": kms,
           "traffic_source": f"profiles/traffic_{workload}.txt: rocprofv3 --pmc FETCH_SIZE and "
                             "WRITE_SIZE of one launch, scaled x2 and x1 by a 1 GiB device copy",
           "model_bytes": r["kernel_bytes"],
           "model_gbs": r["kernel_bytes"] / (kms * 1e-3) / 1e9 if r["kernel_bytes"] else None,
           "model_note": "cost model of the counting path (projection pairs, neighbour-row entries, "
                         "bins at 8 B each); mostly served by LDS, L2 and the Infinity Cache, so not "
                         "an HBM lower bound and not the roofline numerator"}
    valu = measured_valu(r["kernel"], workload)
    probe = _profile_kv("valu_probe.txt")
    if valu and probe.get("valu_wave_instructions_per_s"):
        peak = float(probe["valu_wave_instructions_per_s"])
        rate = valu / (kms * 1e-3)
        out["valu"] = {"issued_per_s": rate, "calibration_per_s": peak, "ratio_to_calibration": rate / peak,
                       "unit": "wave64 VALU instructions/s",
                       "source": f"SQ_INSTS_VALU per launch from profiles/counters_{r['kernel']}_"
                                 f"{workload}.md over this run's kernel time; calibration = the same "
                                 "counter over a VALU-only integer kernel (profiles/valu_probe.txt: "
                                 "one 4-cycle pass per wave64 instruction per SIMD, 0.99 of it reached)",
                       "note": "a ratio near or above 1 means the kernel is VALU-issue bound; above 1, "
                               "some of its counted instructions take less than a full pass"}
    return out


def run_temporal(steps: int, warmup: int, with_cpu: bool) -> dict:
    off, nodes, times = W.temporal_c4()
    delta = W.TEMPORAL_C4_DELTA_US
    g = hm.TemporalGraph(off, nodes, times)
    for _ in range(warmup):
        counts = g.count(delta)
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        counts = g.count(delta)
        kms.append(g.phase_ms("count"))
    t1 = time.perf_counter()
    st = g.stats()
    g.close()
    inst = int(counts.astype(np.int64).sum())
    kbytes = 8 * 2 * st["proj_pairs"] + 8 * st["oriented_row_sum"] + 8 * hm.TEMPORAL_BINS
    return {"ms": (t1 - t0) * 1e3 / steps, "total": inst, "unit": "instances/s",
            "kernel": "k_tcount", "kernel_ms": float(np.mean(kms)), "kernel_bytes": kbytes,
            "extra": {"temporal_hyperedges": int(off.size - 1), "instances": inst,
                      "window_pairs": st["proj_pairs"], "max_window": st["max_window"],
                      "bytes_model_total": hm.algorithmic_bytes_temporal(st)},
            "cpu": cpu_baseline_temporal(off, nodes, times, delta) if with_cpu else None}


def run_static(steps: int, warmup: int, with_cpu: bool) -> dict:
    off, nodes = hm.generate_static(W.STATIC_C1)
    g = hm.Graph(off, nodes)
    for _ in range(warmup):
        counts = g.count_static()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        counts = g.count_static()
        kms.append(g.phase_ms("triangles"))
    t1 = time.perf_counter()
    st = g.stats()
    g.close()
    triples = int(counts.astype(np.int64).sum())
    kbytes = 8 * st["proj_pairs"] + 8 * st["oriented_row_sum"] + 8 * hm.STATIC_BINS
    return {"ms": (t1 - t0) * 1e3 / steps, "total": triples, "unit": "triples/s",
            "kernel": "k_tri", "kernel_ms": float(np.mean(kms)), "kernel_bytes": kbytes,
            "extra": {"hyperedges": int(off.size - 1), "triples": triples,
                      "bytes_model_total": hm.algorithmic_bytes_static(st)},
            "cpu": cpu_baseline_static(off, nodes) if with_cpu else None}


def wedge_identity(counts: np.ndarray, st: dict) -> bool:
    """sum(open) + 3 sum(closed) = sum_a C(d_a, 2), the projected degrees from the projection
    kernels alone (bins 20..25 are the open triples)."""
    c = counts.astype(np.int64)
    closed = int(c.sum() - c[20:26].sum())
    return int(c[20:26].sum()) + 3 * closed == st["wedges"]


def run_static_c3_base(steps: int, warmup: int, with_cpu: bool) -> dict:
    """The full 500k-hyperedge dynamic base by the hub-aware count. The dominant kernel is the
    light-triangle kernel (k_hub_tri_tiled), timed by the count's own HIP events."""
    off, nodes = hm.generate_static(W.DYNAMIC_C3)
    g = hm.Graph(off, nodes)
    g.set_hubs(-1)
    for _ in range(warmup):
        counts = g.count_static()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        counts = g.count_static()
        kms.append(g.hub_stats()["ms_tri_kernel"])
    t1 = time.perf_counter()
    st, hs = g.stats(), g.hub_stats()
    g.close()
    triples = int(counts.astype(np.int64).sum())
    return {"ms": (t1 - t0) * 1e3 / steps, "total": triples, "unit": "triples/s",
            "kernel": "k_hub_tri_tiled", "kernel_ms": float(np.mean(kms)), "kernel_bytes": None,
            "extra": {"hyperedges": int(off.size - 1), "triples": triples, "hub_stats": hs,
                      "wedge_identity": wedge_identity(counts, st),
                      "largest_node_degree": st["max_node_degree"]},
            "cpu": None}


def static_hub_summary() -> dict:
    """Static counts by the hub-aware path beside the triangle path (identical bins): static c1
    (5,000 hyperedges) and a 60,000-hyperedge graph of the 2M-hyperedge graph's family (V = E/2,
    seed 22), each the best of a few calls after one untimed call."""
    out = {}
    for name, p, plain in (("static_c1", W.STATIC_C1, True),
                           ("static_c2_family_60k", W.scaled(W.STATIC_C2, 60_000, 30_000, seed=22), False)):
        off, nodes = hm.generate_static(p)
        rec = {"hyperedges": int(off.size - 1)}
        with hm.Graph(off, nodes) as g:
            if plain:
                base = g.count_static()
                ts = []
                for _ in range(5):
                    t = time.perf_counter()
                    base = g.count_static()
                    ts.append(time.perf_counter() - t)
                rec["triangle_path_ms"] = min(ts) * 1e3
            g.set_hubs(-1)
            c = g.count_static()
            ts = []
            for _ in range(3):
                t = time.perf_counter()
                c = g.count_static()
                ts.append(time.perf_counter() - t)
            st = g.stats()
            rec.update({"hub_path_ms": min(ts) * 1e3, "triples": int(c.astype(np.int64).sum()),
                        "triples_per_s": float(c.astype(np.int64).sum()) / min(ts),
                        "hubs": g.hub_stats()["hubs"], "wedge_identity": wedge_identity(c, st)})
            if plain:
                rec["identical_bins"] = bool(np.array_equal(c, base))
        out[name] = rec
    return out


def run_dynamic(steps: int, warmup: int, with_cpu: bool) -> dict:
    p = W.scaled(W.DYNAMIC_C3, DYN_BASE_EDGES, DYN_BASE_NODES)
    scale = DYN_BASE_EDGES / W.DYNAMIC_C3.num_edges
    n_ins, n_del = round(W.DYNAMIC_C3_INSERTS * scale), round(W.DYNAMIC_C3_DELETES * scale)
    off, nodes = hm.generate_static(p)
    g = hm.Graph(off, nodes)
    # batches are built up front (the recipe depends on the live set only); the host keeps the
    # rows too, for the CPU recount of the final graph
    live = np.arange(DYN_BASE_EDGES, dtype=np.int64)
    rows = [nodes[off[e]:off[e + 1]] for e in range(DYN_BASE_EDGES)]
    batches = []
    for b in range(1, warmup + steps + 1):
        dels, ioff, inodes = W.dynamic_batch(live, b, p, n_ins, n_del)
        batches.append((dels, ioff, inodes))
        nxt = len(rows)
        for x in dels:
            rows[int(x)] = nodes[:0]
        rows.extend(inodes[ioff[k]:ioff[k + 1]] for k in range(ioff.size - 1))
        live = np.concatenate([np.setdiff1d(live, dels), np.arange(nxt, len(rows))])
    deltas = []
    for dels, ioff, inodes in batches[:warmup]:
        deltas.append(g.update(dels, ioff, inodes)[1])
    touched, rows_ms, delta_ms, store_ms, entries = 0, 0.0, 0.0, 0.0, 0
    t0 = time.perf_counter()
    for dels, ioff, inodes in batches[warmup:]:
        deltas.append(g.update(dels, ioff, inodes)[1])
        us = g.update_stats()
        touched += us["removed"] + us["added"]
        rows_ms += us["ms_rows"]
        delta_ms += us["ms_delta"]
        store_ms += us["ms_store"]
        entries += us["row_entries"]
    t1 = time.perf_counter()
    # the same batches as captured updates (one graph launch per update + delta) on a second copy
    gc = hm.Graph(off, nodes)
    gc.set_update_capture(True)
    same_cap = all(np.array_equal(gc.update(*bt)[1], d) for bt, d in zip(batches[:warmup], deltas))
    tc = time.perf_counter()
    for bt, d in zip(batches[warmup:], deltas[warmup:]):
        same_cap = np.array_equal(gc.update(*bt)[1], d) and same_cap
    cap_ms = (time.perf_counter() - tc) * 1e3 / steps
    cap_state = gc.update_capture_state()
    gc.close()
    g.count_static()  # untimed: scratch sized and the triangle form tuned for this shape
    t2 = time.perf_counter()
    total = int(g.count_static().astype(np.int64).sum())  # the recount a batch would need instead
    recount_ms = (time.perf_counter() - t2) * 1e3
    g.close()
    store_full = store_updates_full_c3()
    cur_off = np.concatenate([[0], np.cumsum([r.size for r in rows])]).astype(np.int64)
    cur_nodes = np.concatenate(rows).astype(np.int32)
    # delta kernel bytes: 4 B per scanned packed row entry + 8 B per histogram bin, per launch
    kbytes = (4 * entries + 8 * hm.STATIC_BINS * 2 * steps) / (2 * steps)
    return {"ms": (t1 - t0) * 1e3 / steps, "total": touched, "unit": "triples/s",
            "kernel": "k_delta", "kernel_ms": delta_ms / (2 * steps), "kernel_bytes": kbytes,
            "extra": {"base_hyperedges": DYN_BASE_EDGES, "inserts_per_batch": n_ins,
                      "deletes_per_batch": n_del, "touched_triples": touched,
                      "ms_per_batch": {"rows": rows_ms / steps, "delta": delta_ms / steps,
                                       "store": store_ms / steps},
                      "full_recount_ms": recount_ms, "triples_after": total,
                      "store_updates_full_c3": store_full,
                      "captured_updates": {"ms_per_batch": cap_ms, "deltas_identical": bool(same_cap),
                                           "captures": cap_state["captures"],
                                           "last_ran_captured": bool(cap_state["last_captured"])}},
            "cpu": cpu_baseline_recount(cur_off, cur_nodes) if with_cpu else None}


def store_updates_full_c3() -> dict:
    """The dynamic workload's store update at full size (500,000-hyperedge base, 100 batches of
    7,000 inserts and 3,000 deletes), without the delta: device ms of the store update per batch
    (HIP events) and host wall time per hm_graph_update call."""
    off, nodes = hm.generate_static(W.DYNAMIC_C3)
    live = np.arange(off.size - 1, dtype=np.int64)
    batches = []
    for b in range(1, W.DYNAMIC_C3_BATCHES + 1):
        dels, ioff, inodes = W.dynamic_batch(live, b)
        batches.append((dels, ioff, inodes))
        live = np.concatenate([np.setdiff1d(live, dels.astype(np.int64)),
                               np.arange(off.size - 1 + (b - 1) * W.DYNAMIC_C3_INSERTS,
                                         off.size - 1 + b * W.DYNAMIC_C3_INSERTS, dtype=np.int64)])
    dev, wall = [], []
    with hm.Graph(off, nodes) as g:
        for dels, ioff, inodes in batches:
            t0 = time.perf_counter()
            g.update(dels, ioff, inodes, delta=False)
            wall.append((time.perf_counter() - t0) * 1e3)
            dev.append(g.phase_ms("store"))
        ids_after = g.read_csr()[0].size - 1
    return {"base_hyperedges": int(off.size - 1), "batches": len(batches),
            "ids_after": int(ids_after), "store_ms_per_batch_mean": float(np.mean(dev)),
            "store_ms_per_batch_max": float(np.max(dev)), "update_call_wall_ms_mean": float(np.mean(wall))}


def copy_probe(mib: int) -> None:
    """One device-to-device copy of `mib` MiB through the HIP runtime: the reference kernel whose
    rocprofv3 FETCH_SIZE / WRITE_SIZE readings calibrate the traffic counters."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), n) == 0 and hip.hipMalloc(ctypes.byref(b), n) == 0
    assert hip.hipMemset(a, 1, n) == 0
    assert hip.hipMemcpy(b, a, n, 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": mib << 20}), flush=True)


def valu_probe(iters: int, lanes: int) -> None:
    """The VALU calibration kernel (hm_probe_valu): its SQ_INSTS_VALU reading under rocprofv3 over
    its duration is the issue peak the roofline's valu fraction divides by."""
    ms = hm.probe_valu(iters, lanes)
    print(json.dumps({"valu_probe_iters": iters, "valu_probe_lanes": lanes, "valu_probe_ms": ms}), flush=True)


def serving_summary(slots: int = 4, n_requests: int = 400) -> dict:
    """The serving half of the metric in the default run: one request stream over the reduced
    serving set, plain (one request at a time) against served (hm_server, `slots` concurrent, with
    and without kept counts), saturated throughput and p50/p99 under Poisson arrivals at 70% of
    the plain path's capacity; every result compared request by request."""
    from ipdps26_amd import serving

    graphs = W.serving_graphs()
    reqs = W.serving_requests(graphs, n_requests)
    same = lambda a, b: all(np.array_equal(x, y) for x, y in zip(a["results"], b["results"]))
    keep = ("requests_per_s", "p50_ms", "p99_ms", "mean_ms")
    pick = lambda d: {k: d[k] for k in keep}
    plain = serving.run_plain(graphs, reqs, realtime=False)
    served = serving.run_served(graphs, reqs, realtime=False, slots=slots)
    kept = serving.run_served(graphs, reqs, realtime=False, slots=slots, maintain=True)
    rate = 0.7 * plain["requests_per_s"]
    timed = W.serving_requests(graphs, n_requests, rate_per_s=rate)
    plain_rt = serving.run_plain(graphs, timed, realtime=True)
    served_rt = serving.run_served(graphs, timed, realtime=True, slots=slots)
    kept_rt = serving.run_served(graphs, timed, realtime=True, slots=slots, maintain=True)
    identical = (same(plain, served) and same(plain, kept) and same(plain_rt, served_rt)
                 and same(plain_rt, kept_rt))
    print(f"[serving] plain {plain['requests_per_s']:.0f} req/s, served {served['requests_per_s']:.0f}, "
          f"kept counts {kept['requests_per_s']:.0f}; p99 at {rate:.0f} req/s: plain "
          f"{plain_rt['p99_ms']:.1f} ms, served {served_rt['p99_ms']:.1f}, kept {kept_rt['p99_ms']:.1f}",
          file=sys.stderr, flush=True)
    return {"workload": DESC["serving_c5"], "slots": slots, "requests": n_requests,
            "results_identical_plain_vs_served": bool(identical),
            "saturated": {"plain": pick(plain), "served": pick(served), "served_kept_counts": pick(kept),
                          "speedup_served": served["requests_per_s"] / plain["requests_per_s"],
                          "speedup_kept_counts": kept["requests_per_s"] / plain["requests_per_s"]},
            "poisson_at_70pct_plain_capacity": {"rate_per_s": rate, "plain": pick(plain_rt),
                                                "served": pick(served_rt),
                                                "served_kept_counts": pick(kept_rt)},
            "note": "kept counts: count queries are answered from each graph's counts kept current "
                    "by the exact deltas of its update batches (no launch); the recounting served "
                    "form runs every count query on the device"}


def run_serving(slots: int = 4, n_requests: int = 400, batch_graphs: int = 16) -> dict:
    """Plain (one request at a time) versus served (hm_server, `slots` concurrent) on one request
    stream: throughput with every request queued at t=0, then p50/p99 latency with Poisson
    arrivals at 70% of the plain path's capacity. Results must agree request by request."""
    from ipdps26_amd import serving

    graphs = W.serving_graphs()
    reqs = W.serving_requests(graphs, n_requests)
    def note(msg):
        print(f"[serving] {msg}", file=sys.stderr, flush=True)

    def same_as(a, b):
        return all(np.array_equal(x, y) for x, y in zip(a["results"], b["results"]))

    plain = serving.run_plain(graphs, reqs, realtime=False)
    note(f"plain saturated: {plain['requests_per_s']:.1f} req/s")
    served = serving.run_served(graphs, reqs, realtime=False, slots=slots)
    note(f"served saturated: {served['requests_per_s']:.1f} req/s")
    batched = serving.run_served(graphs, reqs, realtime=False, slots=slots, batch_graphs=batch_graphs)
    note(f"served with batched counts saturated: {batched['requests_per_s']:.1f} req/s")
    kept = serving.run_served(graphs, reqs, realtime=False, slots=slots, maintain=True)
    note(f"served with kept counts saturated: {kept['requests_per_s']:.1f} req/s")
    kept_cap = serving.run_served(graphs, reqs, realtime=False, slots=slots, maintain=True, capture=True)
    note(f"served with kept counts and captured updates saturated: {kept_cap['requests_per_s']:.1f} req/s")
    same = (same_as(plain, served) and same_as(plain, batched) and same_as(plain, kept)
            and same_as(plain, kept_cap))
    rate = 0.7 * plain["requests_per_s"]
    timed = W.serving_requests(graphs, n_requests, rate_per_s=rate)
    plain_rt = serving.run_plain(graphs, timed, realtime=True)
    note(f"plain at {rate:.1f} req/s: p50 {plain_rt['p50_ms']:.1f} ms p99 {plain_rt['p99_ms']:.1f} ms")
    served_rt = serving.run_served(graphs, timed, realtime=True, slots=slots)
    note(f"served at {rate:.1f} req/s: p50 {served_rt['p50_ms']:.1f} ms p99 {served_rt['p99_ms']:.1f} ms")
    batched_rt = serving.run_served(graphs, timed, realtime=True, slots=slots, batch_graphs=batch_graphs)
    note(f"served with batched counts at {rate:.1f} req/s: p50 {batched_rt['p50_ms']:.1f} ms "
         f"p99 {batched_rt['p99_ms']:.1f} ms")
    kept_rt = serving.run_served(graphs, timed, realtime=True, slots=slots, maintain=True)
    note(f"served with kept counts at {rate:.1f} req/s: p50 {kept_rt['p50_ms']:.1f} ms "
         f"p99 {kept_rt['p99_ms']:.1f} ms")
    kept_cap_rt = serving.run_served(graphs, timed, realtime=True, slots=slots, maintain=True, capture=True)
    note(f"served with kept counts and captured updates at {rate:.1f} req/s: "
         f"p50 {kept_cap_rt['p50_ms']:.1f} ms p99 {kept_cap_rt['p99_ms']:.1f} ms")
    same = (same and same_as(plain_rt, served_rt) and same_as(plain_rt, batched_rt)
            and same_as(plain_rt, kept_rt) and same_as(plain_rt, kept_cap_rt))
    # many small graphs: where batched counts pay (launch-bound counts)
    sg = W.serving_graphs(256, lo=100, hi=1000, seed=55)
    sreqs = W.serving_requests(sg, 2000, seed=5500)
    s_plain = serving.run_plain(sg, sreqs, realtime=False)
    s_served = serving.run_served(sg, sreqs, realtime=False, slots=slots)
    s_batched = serving.run_served(sg, sreqs, realtime=False, slots=slots, batch_graphs=64)
    note(f"small graphs saturated: plain {s_plain['requests_per_s']:.0f} req/s, served "
         f"{s_served['requests_per_s']:.0f}, served with batched counts {s_batched['requests_per_s']:.0f}")
    same = same and same_as(s_plain, s_served) and same_as(s_plain, s_batched)
    # the full configuration's own graphs up to 30k hyperedges, at their sizes
    band = W.serving_c5_band()
    hub_min, budget = W.SERVING_HUB_MIN_EDGES, 160 << 30
    breqs = W.serving_requests(band, 48, seed=5600)
    b_plain = serving.run_plain(band, breqs, realtime=False, hub_min_edges=hub_min)
    note(f"c5 graphs up to {W.SERVING_BAND_MAX} hyperedges, plain saturated: {b_plain['requests_per_s']:.2f} req/s")
    b_served = serving.run_served(band, breqs, realtime=False, slots=slots, budget_bytes=budget,
                                  hub_min_edges=hub_min)
    b_kept = serving.run_served(band, breqs, realtime=False, slots=slots, maintain=True, budget_bytes=budget,
                                hub_min_edges=hub_min)
    note(f"c5 graphs up to {W.SERVING_BAND_MAX} hyperedges saturated: plain {b_plain['requests_per_s']:.2f} "
         f"req/s, served {b_served['requests_per_s']:.2f}, kept counts {b_kept['requests_per_s']:.2f}")
    brate = 0.7 * b_plain["requests_per_s"]
    btimed = W.serving_requests(band, 48, seed=5600, rate_per_s=brate)
    b_plain_rt = serving.run_plain(band, btimed, realtime=True, hub_min_edges=hub_min)
    b_kept_rt = serving.run_served(band, btimed, realtime=True, slots=slots, maintain=True, budget_bytes=budget,
                                   hub_min_edges=hub_min)
    note(f"c5 graphs up to {W.SERVING_BAND_MAX} hyperedges at {brate:.2f} req/s: p50/p99 plain "
         f"{b_plain_rt['p50_ms']:.0f}/{b_plain_rt['p99_ms']:.0f} ms, kept counts "
         f"{b_kept_rt['p50_ms']:.0f}/{b_kept_rt['p99_ms']:.0f} ms")
    same = (same and same_as(b_plain, b_served) and same_as(b_plain, b_kept)
            and same_as(b_plain_rt, b_kept_rt))
    triples = serving.count_triples(reqs, served["results"])
    drop = ("results",)

    def strip(d):
        return {k: v for k, v in d.items() if k not in drop}

    return {"ms": served["makespan_s"] * 1e3, "total": triples, "unit": "triples/s",
            "extra": {"requests": n_requests, "slots": slots, "batch_graphs": batch_graphs,
                      "results_identical_plain_vs_served": bool(same),
                      "value_is": "triples classified by the recounting served run / its makespan",
                      "saturated": {"plain": strip(plain), "served": strip(served),
                                    "served_batched_counts": strip(batched),
                                    "served_kept_counts": strip(kept),
                                    "served_kept_counts_captured_updates": strip(kept_cap),
                                    "speedup": served["requests_per_s"] / plain["requests_per_s"],
                                    "speedup_batched_counts": batched["requests_per_s"] / plain["requests_per_s"],
                                    "speedup_kept_counts": kept["requests_per_s"] / plain["requests_per_s"],
                                    "speedup_kept_counts_captured_updates":
                                        kept_cap["requests_per_s"] / plain["requests_per_s"]},
                      "poisson_at_70pct_plain_capacity": {
                          "rate_per_s": rate, "plain": strip(plain_rt), "served": strip(served_rt),
                          "served_batched_counts": strip(batched_rt),
                          "served_kept_counts": strip(kept_rt),
                          "served_kept_counts_captured_updates": strip(kept_cap_rt)},
                      "small_graphs_saturated": {
                          "graphs": "256 graphs of the static_c2 family, E log-uniform on [100, 1000], "
                                    "seed 55; 2,000 requests (seed 5500), all queued at t=0",
                          "plain": strip(s_plain), "served": strip(s_served),
                          "served_batched_counts_64": strip(s_batched),
                          "speedup_batched_counts": s_batched["requests_per_s"] / s_plain["requests_per_s"]},
                      "c5_graphs_at_size": {
                          "graphs": f"the {len(band)} graphs of the full serving configuration (64 graphs, "
                                    f"E log-uniform on [10k, 1M], seed 5) with at most {W.SERVING_BAND_MAX} "
                                    "hyperedges, each at its own size; 48 requests (seed 5600); graphs of "
                                    f"{hub_min} hyperedges or more counted and updated hub-aware in every form",
                          "hyperedges": [int(off.size - 1) for off, _ in band],
                          "saturated": {"plain": strip(b_plain), "served": strip(b_served),
                                        "served_kept_counts": strip(b_kept),
                                        "speedup_kept_counts": b_kept["requests_per_s"] / b_plain["requests_per_s"]},
                          "poisson_at_70pct_plain_capacity": {"rate_per_s": brate, "plain": strip(b_plain_rt),
                                                              "served_kept_counts": strip(b_kept_rt)}}}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (counter calibration)")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="temporal_c4", choices=sorted(DESC))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the shortened serving stream of the default run")
    ap.add_argument("--valu-probe", type=int, default=0, metavar="ITERS",
                    help="only run the VALU calibration kernel (counter calibration)")
    ap.add_argument("--valu-lanes", type=int, default=64,
                    help="active lanes per wave of the VALU calibration kernel (1..64)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return 0
    if args.valu_probe:
        hm.init(0)
        valu_probe(args.valu_probe, args.valu_lanes)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the engine has no "
              "multi-GPU path", file=sys.stderr)
        return 2
    steps = args.steps if args.steps is not None else DEFAULT_STEPS[args.workload]
    hm.init(0)
    if args.workload == "serving_c5":
        r = run_serving()
        line = {"metric": "exact h-motif triples classified/sec; p50/p99 latency, served vs plain",
                "value": r["total"] / (r["ms"] * 1e-3), "unit": r["unit"], "n_gpus": 1,
                "steps": 1, "warmup": 0, "ms_per_step": r["ms"], "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "u64",
                "data": "synthetic (seeded generator)",
                "config": {"workload": DESC[args.workload], **r["extra"]},
                "roofline": None, "cpu_baseline": None}
        print(json.dumps(line), flush=True)
        return 0
    run = {"temporal_c4": run_temporal, "static_c1": run_static, "dynamic_c3": run_dynamic,
           "static_c3_base": run_static_c3_base}[args.workload]
    warmup = 0 if args.workload == "static_c3_base" and args.steps is None else args.warmup
    r = run(steps, warmup, not args.no_cpu)
    config = {"workload": DESC[args.workload], **r["extra"]}
    if args.workload == "temporal_c4":
        config["why_this_workload"] = WHY_DEFAULT
        if not args.no_serving:
            config["serving"] = serving_summary()
            config["static_hub"] = static_hub_summary()
    if r["cpu"] is not None:
        r["cpu"]["host"] = host_cpu()
    line = {
        "metric": "exact h-motif triples classified/sec",
        "value": r["total"] / (r["ms"] * 1e-3),
        "unit": r["unit"],
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": r["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded generator)",
        "config": config,
        "roofline": roofline(args.workload, r),
        "cpu_baseline": r["cpu"],
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
