"""The exact 30 bins of the 2M-hyperedge static configuration (1M nodes, sizes ~k^-2.2 on [2, 50],
Zipf s=1.1, seed 2), counted hub-aware in parts over several GPU runs (hm_graph_count_hub_part),
since its light triangles take longer than one run:

  python tests/golden/make_static_c2.py --hubs H --other --begin 0 --seconds S --out DIR
      the class, pair-type and light-wedge parts with the projection statistics, then the light
      triangles of light ranks 0, 1, ... until S seconds have passed;
  python tests/golden/make_static_c2.py --hubs H --begin R --seconds S --out DIR
      the light triangles from rank R on, until S seconds have passed (R = the previous part's end);
  python tests/golden/make_static_c2.py --combine DIR
      for every hub count with parts in DIR: checks that exactly one part holds the other parts and
      that the ranges tile [0, E), sums the bins (mod 2^64) and checks sum(open) + 3 sum(closed) =
      sum_a C(d_a, 2) against the projected degrees; the hub counts must agree bin for bin; writes
      static_c2_hub.json beside this script.

Each part writes DIR/c2_h<H>_part_<begin>.json. Needs an MI355X for the parts; no CPU oracle finishes
this graph, so the result is a regression vector held to the wedge identity, not an oracle pin.
Run from the repository root."""
import argparse
import glob
import json
import os
import sys
import time

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
import ipdps26_amd as hm  # noqa: E402
from ipdps26_amd import workloads as W  # noqa: E402

HUBS = 23  # the hub count of the fixture's bins
OPEN = range(20, 26)


def run_part(args) -> int:
    os.environ["HM_VERBOSE"] = "1"  # the count reports each phase and each launch of centres
    hm.init(0)
    t = time.perf_counter()
    off, nodes = hm.generate_static(W.STATIC_C2)
    E = off.size - 1
    print(f"graph: {E} hyperedges, {nodes.size} incidences, generated in {time.perf_counter() - t:.1f} s",
          flush=True)
    with hm.Graph(off, nodes) as g:
        t = time.perf_counter()
        bins, lt, done = g.count_hub_part(args.hubs, args.begin, -1, other_parts=args.other,
                                          seconds=args.seconds, with_done=True)
        secs = time.perf_counter() - t
        rec = {"hubs": args.hubs, "hyperedges": int(E), "begin": args.begin, "end": done,
               "other_parts": bool(args.other), "bins": [int(x) for x in bins], "light_triangles": lt,
               "seconds": secs, "hub_stats": g.hub_stats()}
        if args.other:
            st = g.stats()
            rec["projection"] = {k: int(st[k]) for k in ("wedges", "proj_pairs", "max_proj_degree",
                                                         "max_node_degree")}
    os.makedirs(args.out, exist_ok=True)
    with open(os.path.join(args.out, f"c2_h{args.hubs}_part_{args.begin}.json"), "w") as f:
        json.dump(rec, f)
    print(f"part from rank {args.begin}: ranks up to {done} of {E} in {secs:.1f} s, {lt} light triangles, "
          f"other parts: {bool(args.other)}", flush=True)
    return 0


def combine_hubs(parts):
    """Bins, other-parts record and light triangles of one hub count's parts, after the checks."""
    parts = sorted(parts, key=lambda r: r["begin"])
    E = parts[0]["hyperedges"]
    assert all(r["hyperedges"] == E for r in parts)
    assert sum(r["other_parts"] for r in parts) == 1, "the other parts must be counted exactly once"
    assert parts[0]["begin"] == 0 and parts[-1]["end"] == E, "ranges must start at 0 and end at E"
    assert all(a["end"] == b["begin"] for a, b in zip(parts, parts[1:])), "ranges must be contiguous"
    bins = [sum(r["bins"][k] for r in parts) % (1 << 64) for k in range(30)]
    other = next(r for r in parts if r["other_parts"])
    wedges = other["projection"]["wedges"]
    n_open = sum(bins[k] for k in OPEN)
    assert n_open + 3 * (sum(bins) - n_open) == wedges, "wedge identity"
    return parts, bins, other, sum(r["light_triangles"] for r in parts)


def combine(args) -> int:
    records = [json.load(open(p)) for p in glob.glob(os.path.join(args.combine, "c2_h*_part_*.json"))]
    assert records, "no parts"
    by_hubs = {}
    for r in records:
        by_hubs.setdefault(r["hubs"], []).append(r)
    runs = {h: combine_hubs(ps) for h, ps in by_hubs.items()}
    assert HUBS in runs, f"no parts with {HUBS} hubs"
    parts, bins, other, lt = runs[HUBS]
    for h, (_, b, o, _) in runs.items():
        print(f"{h} hubs: {len(by_hubs[h])} parts, {sum(b)} triples, wedge identity holds", flush=True)
        assert b == bins, f"{h} and {HUBS} hubs disagree"
        assert o["projection"] == other["projection"]
    p = W.STATIC_C2
    out = {"name": "static_c2",
           "params": {k: getattr(p, k) for k in ("num_nodes", "num_edges", "alpha", "kmin", "kmax", "zipf_s",
                                                 "seed")},
           "hubs": HUBS, "counts30": bins, "triples": sum(bins), "wedges": other["projection"]["wedges"],
           "projection": other["projection"], "light_triangles": lt,
           "parts": [{"begin": r["begin"], "end": r["end"], "other_parts": r["other_parts"],
                      "light_triangles": r["light_triangles"], "seconds": round(r["seconds"], 1)} for r in parts],
           "hub_counts_agreeing": sorted(runs, reverse=True),
           "light_triangles_by_hubs": {str(h): runs[h][3] for h in sorted(runs, reverse=True)},
           "source": "hub-aware GPU counts in parts (hm_graph_count_hub_part) on one MI355X, identical bin for "
                     "bin with every hub count in hub_counts_agreeing; sum(open) + 3 sum(closed) = "
                     "sum_a C(d_a, 2) holds against the projected degrees. A regression vector, not an oracle "
                     "pin (no CPU oracle finishes this graph)."}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "static_c2_hub.json"), "w") as f:
        json.dump(out, f)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--hubs", type=int, default=HUBS, help="hub count (1..23)")
    ap.add_argument("--other", action="store_true", help="also count the class, pair-type and light-wedge parts")
    ap.add_argument("--begin", type=int, default=0, help="first light rank of the light-triangle part")
    ap.add_argument("--seconds", type=float, default=0.0, help="time limit of the light-triangle part")
    ap.add_argument("--out", help="directory for the part's JSON file")
    ap.add_argument("--combine", metavar="DIR", help="combine the parts in DIR")
    args = ap.parse_args()
    if not args.combine and not args.out:
        ap.error("--out is required for a part")
    return combine(args) if args.combine else run_part(args)


if __name__ == "__main__":
    sys.exit(main())
