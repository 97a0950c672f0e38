"""Writes full-size count vectors of two workloads from the CPU oracle (build container only).

    python tests/golden/make_oracle_full.py [--static] [--temporal] [--workers N] [--shards S]

static    tests/golden/static_c1_30bins.json: all 30 bins of the 5,000-hyperedge static workload
          from oracle_static_mochy (the restatement of MoCHy-E,
          MoCHy-master/main_exact_par.cpp:74-173). MoCHy prints 26 of the 30 bins
          (main_exact.cpp:133-137); the 26 printed ones are checked here against the reference
          MoCHy golden of the same input (static_golden.json, "static_c1"), so the file adds the
          hidden bins 0, 1, 4 and 6.
temporal  tests/golden/temporal_c4_full.json: the 96 THyMe bins of the full temporal workload
          (200,000 temporal hyperedges, delta = 3,600 s) from oracle_thyme_temporal_range (the
          restatement of THyMe+, THyMe-main/code/main_thymeP.cpp:219-367) over S anchor ranges in
          a pool of N processes; the ranges sum to the whole count (tests/test_oracle.py checks the
          range form against the reference goldens and the brute force).

Inputs come from libhmotif's seeded generator; each file records the generator parameters and
the SHA-256 of the exact input text (MoCHy CSV / THyMe TSV).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import multiprocessing as mp
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle_lib as O  # noqa: E402
import ipdps26_amd as hm  # noqa: E402
from ipdps26_amd import workloads as W  # noqa: E402


def sha256_file(path: str) -> str:
    h = hashlib.sha256()
    with open(path, "rb") as f:
        h.update(f.read())
    return h.hexdigest()


def write_json(path: str, obj) -> None:
    with open(path, "w") as f:
        json.dump(obj, f, indent=1)
        f.write("\n")


def make_static(threads: int) -> None:
    p = W.STATIC_C1
    off, nodes = hm.generate_static(p)
    with tempfile.TemporaryDirectory() as td:
        csv = os.path.join(td, "c1.csv")
        W.write_mochy_csv(csv, off, nodes)
        sha = sha256_file(csv)
    t0 = time.time()
    bins = O.static_mochy(off, nodes, threads).tolist()
    secs = time.time() - t0
    ref = [g for g in json.load(open(os.path.join(HERE, "static_golden.json"))) if g["name"] == "static_c1"]
    printed = [bins[b] for b in [2, 3, 5] + list(range(7, 30))]
    if ref:
        assert ref[0]["sha256_csv"] == sha, "input differs from the reference golden's"
        assert ref[0]["printed_motifs"] == printed, "oracle disagrees with the reference MoCHy"
    write_json(os.path.join(HERE, "static_c1_30bins.json"), {
        "name": "static_c1_30bins", "params": p.__dict__, "sha256_csv": sha, "counts30": bins,
        "hidden_bins": {str(b): bins[b] for b in (0, 1, 4, 6)},
        "source": "oracle_static_mochy (restatement of MoCHy-E); printed bins equal the reference "
                  "MoCHy golden static_c1", "oracle_seconds": round(secs, 2), "threads": threads})
    print(f"static c1: {sum(bins)} triples, hidden {[bins[b] for b in (0, 1, 4, 6)]}, {secs:.1f} s",
          flush=True)


_G = None


def _init() -> None:
    global _G
    _G = W.temporal_c4()


def _shard(rng):
    lo, hi = rng
    off, nodes, times = _G
    t0 = time.time()
    c = O.thyme_temporal_range(off, nodes, times, W.TEMPORAL_C4_DELTA_US, lo, hi)
    return lo, hi, c, time.time() - t0


def make_temporal(workers: int, shards: int) -> None:
    off, nodes, times = W.temporal_c4()
    E = off.size - 1
    with tempfile.TemporaryDirectory() as td:
        tsv = os.path.join(td, "c4.tsv")
        W.write_thyme_tsv(tsv, off, nodes, times)
        sha = sha256_file(tsv)
    cuts = np.linspace(0, E, shards + 1).round().astype(int).tolist()
    # most-work-last ranges first would need a work model; the pool's unordered map balances
    ranges = list(zip(cuts[:-1], cuts[1:]))
    total = np.zeros(96, dtype=np.int64)
    cpu, done, t0 = 0.0, 0, time.time()
    with mp.Pool(workers, initializer=_init) as pool:
        for lo, hi, c, s in pool.imap_unordered(_shard, ranges):
            total += c
            cpu += s
            done += 1
            print(f"shard [{lo}, {hi}) {int(c.sum())} instances {s:.1f} s; {done}/{len(ranges)} "
                  f"after {time.time() - t0:.0f} s", flush=True)
    p = W.TEMPORAL_C4
    write_json(os.path.join(HERE, "temporal_c4_full.json"), {
        "name": "temporal_c4_full", "params": p.__dict__, "delta": W.TEMPORAL_C4_DELTA_US,
        "timestamps": "generate_timestamps(num_edges, seed)", "sha256_tsv": sha,
        "counts96": total.tolist(), "instances": int(total.sum()),
        "source": "oracle_thyme_temporal_range (restatement of THyMe+) summed over anchor ranges",
        "shards": len(ranges), "workers": workers, "oracle_cpu_seconds": round(cpu, 1),
        "wall_seconds": round(time.time() - t0, 1)})
    print(f"temporal c4: {int(total.sum())} instances, {cpu:.0f} CPU s", flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--static", action="store_true")
    ap.add_argument("--temporal", action="store_true")
    ap.add_argument("--workers", type=int, default=max(1, (os.cpu_count() or 2) - 1))
    ap.add_argument("--shards", type=int, default=800)
    a = ap.parse_args()
    if not (a.static or a.temporal):
        a.static = a.temporal = True
    if a.static:
        make_static(a.workers)
    if a.temporal:
        make_temporal(a.workers, a.shards)
    return 0


if __name__ == "__main__":
    sys.exit(main())
