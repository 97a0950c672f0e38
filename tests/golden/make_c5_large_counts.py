"""The exact 30 bins of the serving configuration's graphs above the served band (64 graphs of the
2M-hyperedge graph's family, hyperedge counts log-uniform on [10k, 1M], seed 5: the 26 with more
than 130,000 hyperedges, 191k to 996k), each counted once at its own size by the hub-aware count
with the automatic hub count. Each count is checked against sum(open) + 3 sum(closed) =
sum_a C(d_a, 2) from the projection kernels. Graphs are counted in ascending size from --first on;
no new graph starts after --seconds, so a later run continues where one stopped:

  python tests/golden/make_c5_large_counts.py --first K --seconds S --out DIR
  python tests/golden/make_c5_large_counts.py --combine DIR     (writes serving_c5_large_counts.json)

Needs an MI355X for the counts; a regression vector held to the wedge identity, not an oracle pin.
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

OPEN = range(20, 26)


def large_graphs():
    """(index in the configuration, hyperedges) of the graphs above the served band, ascending."""
    sizes = W.serving_c5_sizes(W.SERVING_C5_GRAPHS, 10_000, 1_000_000, 5)
    return sorted(((i, e) for i, e in enumerate(sizes) if e > W.SERVING_BAND_MAX), key=lambda t: t[1])


def run(args) -> int:
    os.environ["HM_VERBOSE"] = "1"  # each count reports its phases and launches
    hm.init(0)
    os.makedirs(args.out, exist_ok=True)
    todo = large_graphs()[args.first:]
    t_start = time.perf_counter()
    for k, (i, e) in enumerate(todo):
        if time.perf_counter() - t_start > args.seconds:
            print(f"time is up before graph {args.first + k} ({e} hyperedges)", flush=True)
            break
        off, nodes = W._serving_graph(e, i, 5)
        with hm.Graph(off, nodes) as g:
            g.set_hubs(-1)
            t = time.perf_counter()
            c = [int(x) for x in g.count_static()]
            secs = time.perf_counter() - t
            st, hs = g.stats(), g.hub_stats()
        n_open = sum(c[b] for b in OPEN)
        ok = n_open + 3 * (sum(c) - n_open) == st["wedges"]
        rec = {"rank": args.first + k, "graph": i, "hyperedges": e, "counts30": c, "triples": sum(c),
               "wedges": st["wedges"], "wedge_identity": ok, "seconds": secs, "hubs": hs["hubs"],
               "light_triangles": hs["light_triangles"], "max_node_degree": st["max_node_degree"]}
        with open(os.path.join(args.out, f"c5_graph_{i}.json"), "w") as f:
            json.dump(rec, f)
        print(f"graph {i} ({e} hyperedges): {sum(c)} triples in {secs:.1f} s, {hs['hubs']} hubs, "
              f"wedge identity {ok}", flush=True)
        if not ok:
            return 1
    return 0


def combine(args) -> int:
    recs = {r["graph"]: r for r in (json.load(open(p)) for p in glob.glob(os.path.join(args.combine, "*.json")))}
    want = large_graphs()
    missing = [i for i, _ in want if i not in recs]
    assert not missing, f"graphs not counted: {missing}"
    assert all(recs[i]["wedge_identity"] and recs[i]["hyperedges"] == e for i, e in want)
    out = {"name": "serving_c5_large",
           "graphs": "the serving configuration's graphs with more than "
                     f"{W.SERVING_BAND_MAX} hyperedges (64 graphs of the 2M-hyperedge graph's family, "
                     "E log-uniform on [10k, 1M], V = E/2, seed 5000 + index), each at its own size",
           "counts": [recs[i] for i, _ in want],
           "source": "hub-aware GPU counts (automatic hub count) on one MI355X, each satisfying "
                     "sum(open) + 3 sum(closed) = sum_a C(d_a, 2) against the projected degrees. A "
                     "regression vector, not an oracle pin (no CPU oracle finishes these graphs)."}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "serving_c5_large_counts.json"), "w") as f:
        json.dump(out, f)
    print(f"{len(want)} graphs, {sum(r['seconds'] for r in out['counts']):.0f} s of counting", flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--first", type=int, default=0, help="first graph (in ascending size) to count")
    ap.add_argument("--seconds", type=float, default=900.0, help="no new graph starts after this long")
    ap.add_argument("--out", help="directory for the counts' JSON files")
    ap.add_argument("--combine", metavar="DIR")
    args = ap.parse_args()
    if not args.combine and not args.out:
        ap.error("--out is required for counting")
    return combine(args) if args.combine else run(args)


if __name__ == "__main__":
    sys.exit(main())
