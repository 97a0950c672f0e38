"""Writes the golden count vectors under tests/golden/ by running the REFERENCE programs.

Run in the build container (needs /root/reference and `make -C oracle ref`):
    python tests/golden/make_golden.py [--skip-c1]

Each vector comes from the reference's own binaries, compiled from their sources by
oracle/Makefile into oracle/_ref/:
  static   MoCHy main_exact_par (8 threads): the 26 printed motifs "motif n: c"
           (MoCHy-master/main_exact_par.cpp:167-173)
  temporal THyMe+ main_thymeP (8 threads): result file lines "i<TAB>count", i = 1..96
           (THyMe-main/code/main_thymeP.cpp:384-389); only inputs with distinct timestamps are
           used, because THyMe breaks ties at random (THyMe-main/code/motif_id.cpp:68-78).
Inputs are either fixtures committed here (dblp_graph.txt and email-Enron-full.txt are data files
of the reference) or outputs of libhmotif's seeded generator, recorded by their parameters and the
SHA-256 of the exact file handed to the reference program.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)

import ipdps26_amd as hm  # noqa: E402
from ipdps26_amd import workloads as W  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")


def sha256(path: str) -> str:
    h = hashlib.sha256()
    with open(path, "rb") as f:
        h.update(f.read())
    return h.hexdigest()


def run_mochy(csv: str, threads: int = 8) -> tuple[list[int], float]:
    out = subprocess.run([os.path.join(REF, "mochy_exact_par"), str(threads), csv],
                         capture_output=True, text=True, check=True).stdout
    counts = [int(m.group(2)) for m in re.finditer(r"^motif (\d+): (\d+)$", out, re.M)]
    assert len(counts) == 26, out[-2000:]
    rt = float(re.search(r"TOTAL_RUNTIME_SECONDS ([0-9.eE+-]+)", out).group(1))
    return counts, rt


def run_thymep(tsv: str, delta: int, threads: int = 8) -> tuple[list[int], float]:
    with tempfile.TemporaryDirectory() as td:
        res = os.path.join(td, "res.txt")
        subprocess.run([os.path.join(REF, "thymeP"), "x", str(delta), str(threads), tsv, res],
                       capture_output=True, text=True, check=True)
        lines = open(res).read().split("\n")
    rt = float(lines[0])
    counts = [int(l.split("\t")[1]) for l in lines[1:97]]
    assert len(counts) == 96
    return counts, rt


def gen_entry(name: str, p: hm.GenParams, td: str) -> dict:
    off, nodes = hm.generate_static(p)
    csv = os.path.join(td, name + ".csv")
    W.write_mochy_csv(csv, off, nodes)
    counts, rt = run_mochy(csv)
    print(f"{name}: E={p.num_edges} mochy {rt:.2f}s", flush=True)
    return {"name": name, "kind": "generated", "params": p.__dict__, "sha256_csv": sha256(csv),
            "printed_motifs": counts, "reference": "MoCHy main_exact_par", "ref_seconds": rt}


def _load(path: str) -> list:
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return []


def _save(path: str, entries: list) -> None:
    with open(path, "w") as f:
        json.dump(entries, f, indent=1)


def main() -> None:
    """Computes every vector not yet present in the JSON files (each is saved once computed)."""
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-c1", action="store_true", help="skip the ~1 min full-size c1 run")
    args = ap.parse_args()
    spath = os.path.join(HERE, "static_golden.json")
    tpath = os.path.join(HERE, "temporal_golden.json")
    static, temporal = _load(spath), _load(tpath)
    have_s = {g["name"] for g in static}
    have_t = {g["name"] for g in temporal}
    with tempfile.TemporaryDirectory() as td:
        if "dblp" not in have_s:
            dblp = os.path.join(HERE, "dblp_graph.txt")
            counts, rt = run_mochy(dblp)
            static.append({"name": "dblp", "kind": "fixture", "file": "dblp_graph.txt",
                           "sha256_csv": sha256(dblp), "printed_motifs": counts,
                           "reference": "MoCHy main_exact_par", "ref_seconds": rt})
            _save(spath, static)
        gens = [
            ("tiny_dense_e60", hm.GenParams(num_nodes=12, num_edges=60, alpha=1.5, kmin=1, kmax=6,
                                            zipf_s=0.8, seed=11)),
            ("c1_family_e800", W.scaled(W.STATIC_C1, 800, 400)),
            ("c2_family_e3000", W.scaled(W.STATIC_C2, 3000, 1500)),
            ("c3_family_e6000", W.scaled(W.DYNAMIC_C3, 6000, 2400)),
            ("c2_family_e8000", W.scaled(W.STATIC_C2, 8000, 4000)),
            ("c5_family_e10000", W.scaled(W.STATIC_C2, 10000, 5000, seed=5)),
        ]
        if not args.skip_c1:
            gens.append(("static_c1", W.STATIC_C1))
        for name, p in gens:
            if name not in have_s:
                static.append(gen_entry(name, p, td))
                _save(spath, static)

        # temporal: email-Enron re-timed to 1000 * line number (distinct), two windows
        enron = os.path.join(HERE, "email-Enron-full.txt")
        off, nodes, _ = W.read_thyme_tsv(enron)
        times = (np.arange(off.size - 1, dtype=np.int64) + 1) * 1000
        tsv = os.path.join(td, "enron_distinct.tsv")
        W.write_thyme_tsv(tsv, off, nodes, times)
        for delta in (100_000, 1_000_000):
            name = f"enron_distinct_d{delta}"
            if name in have_t:
                continue
            counts, rt = run_thymep(tsv, delta)
            print(f"enron delta={delta}: thymeP {rt:.2f}s", flush=True)
            temporal.append({"name": name, "kind": "fixture",
                             "file": "email-Enron-full.txt", "retime": "1000*line_number",
                             "delta": delta, "sha256_tsv": sha256(tsv), "counts96": counts,
                             "reference": "THyMe+ main_thymeP", "ref_seconds": rt})
            _save(tpath, temporal)
        # reduced instances of the temporal workload (THyMe+ enumerates serially, so the windows
        # are kept to a few hundred hyperedges)
        for name, ne, nv, delta in [("c4_family_e3000", 3000, 1500, 600_000_000),
                                    ("c4_family_e12000", 12000, 6000, 120_000_000)]:
            if name in have_t:
                continue
            off, nodes, times = W.temporal_c4(ne, nv)
            tsv = os.path.join(td, name + ".tsv")
            W.write_thyme_tsv(tsv, off, nodes, times)
            counts, rt = run_thymep(tsv, delta)
            print(f"{name}: thymeP {rt:.2f}s", flush=True)
            temporal.append({"name": name, "kind": "generated",
                             "params": W.scaled(W.TEMPORAL_C4, ne, nv).__dict__, "delta": delta,
                             "sha256_tsv": sha256(tsv), "counts96": counts,
                             "reference": "THyMe+ main_thymeP", "ref_seconds": rt})
            _save(tpath, temporal)


if __name__ == "__main__":
    main()
