"""Writes workloads_manifest.json: SHA-256 of every named workload as generated by libhmotif's
seeded generator (offsets int64 LE bytes, then nodes int32 LE bytes, then timestamps int64 LE
bytes for temporal workloads, then for the dynamic workload the first 3 batches' deleted ids and
inserted CSR). tests/test_manifest.py regenerates them and compares, so any change of the
generator (or of a platform's arithmetic) shows up as a changed input, not as changed counts.

    python tests/golden/make_manifest.py
"""
import hashlib
import json
import os
import sys

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from ipdps26_amd import generate_static  # noqa: E402
from ipdps26_amd import workloads as W  # noqa: E402


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<"), copy=False).tobytes())
    return h.hexdigest()


def entries():
    out = {}
    for name, p in [("static_c1", W.STATIC_C1), ("static_c2", W.STATIC_C2),
                    ("dynamic_c3_base", W.DYNAMIC_C3)]:
        off, nodes = generate_static(p)
        out[name] = {"params": p.__dict__, "num_incidences": int(nodes.size), "sha256": digest(off, nodes)}
    off, nodes, times = W.temporal_c4()
    out["temporal_c4"] = {"params": W.TEMPORAL_C4.__dict__, "num_incidences": int(nodes.size),
                          "sha256": digest(off, nodes, times)}
    live = np.arange(W.DYNAMIC_C3.num_edges, dtype=np.int64)
    parts = []
    nxt = W.DYNAMIC_C3.num_edges
    for b in range(1, 4):
        dels, ioff, inodes = W.dynamic_batch(live, b)
        parts += [dels, ioff, inodes]
        live = np.concatenate([np.setdiff1d(live, dels.astype(np.int64)),
                               np.arange(nxt, nxt + ioff.size - 1, dtype=np.int64)])
        nxt += ioff.size - 1
    out["dynamic_c3_batches_1_3"] = {"sha256": digest(*parts)}
    out["serving_c5_sizes"] = {"sizes": W.serving_c5_sizes()}
    return out


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(__file__), "workloads_manifest.json")
    with open(path, "w") as f:
        json.dump(entries(), f, indent=1)
    print("wrote", path)
