"""Writes static_c3_base_hub.json: the 30 bins of the 500,000-hyperedge dynamic base graph
(dynamic workload, seed 3) from the hub-aware GPU count with 12 and with 10 hubs, which must agree
bin for bin and satisfy sum(open) + 3 sum(closed) = sum_a C(d_a, 2). Needs an MI355X; no CPU
oracle finishes this graph. Run from the repository root: python tests/golden/make_hub_c3.py"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
import ipdps26_amd as hm  # noqa: E402
from ipdps26_amd import workloads as W  # noqa: E402


def main():
    hm.init(0)
    off, nodes = hm.generate_static(W.DYNAMIC_C3)
    runs = {}
    for h in (12, 10):
        with hm.Graph(off, nodes) as g:
            g.set_hubs(h)
            c = g.count_static().astype(np.int64)
            st, hs = g.stats(), g.hub_stats()
        closed = int(c.sum() - c[20:26].sum())
        assert int(c[20:26].sum()) + 3 * closed == st["wedges"], h
        runs[h] = (c.tolist(), st, hs)
        print(f"hubs={h}: {int(c.sum())} triples, {hs}", flush=True)
    assert runs[12][0] == runs[10][0]
    digest = hashlib.sha256()
    for e in range(off.size - 1):
        digest.update((",".join(str(int(v)) for v in nodes[off[e]:off[e + 1]]) + "\n").encode())
    p = W.DYNAMIC_C3
    out = {"name": "static_c3_base", "params": {k: getattr(p, k) for k in
                                                ("num_nodes", "num_edges", "alpha", "kmin", "kmax", "zipf_s", "seed")},
           "sha256_csv": digest.hexdigest(), "counts30": runs[12][0], "triples": int(sum(runs[12][0])),
           "wedges": runs[12][1]["wedges"],
           "source": "hub-aware GPU count with 12 and with 10 hubs (identical bins, wedge identity holds); "
                     "a regression vector, not an oracle pin"}
    with open(os.path.join(os.path.dirname(__file__), "static_c3_base_hub.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
