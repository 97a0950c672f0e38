"""Named synthetic workloads (fixed seeds) and file formats of the reference programs.

Workloads (sizes and seeds are part of the definition; the generator is libhmotif's own):
  static_c1     2,000 nodes, 5,000 hyperedges, sizes ~ k^-2.5 on [2, 25], Zipf s = 1.2, seed 1
  static_c2     1,000,000 nodes, 2,000,000 hyperedges, sizes ~ k^-2.2 on [2, 50], Zipf s = 1.1, seed 2
  dynamic_c3    500,000 hyperedges over 200,000 nodes, sizes ~ k^-2.2 on [2, 30], Zipf s = 1.1, seed 3;
                batch b (1..100): 7,000 fresh hyperedges (seed 3000 + b) and 3,000 deletes of live
                hyperedges chosen uniformly (seed 3000 + b)
  temporal_c4   200,000 hyperedges over 100,000 nodes, sizes ~ k^-2.2 on [2, 20], Zipf s = 1.2, seed 4;
                int64 microsecond timestamps: singleton gaps Exp(1 s), 1,200 bursts of 50 with
                Exp(10 ms) gaps; window delta = 3,600 s
  serving_c5    64 graphs, hyperedge counts log-uniform on [10k, 1M] (static_c2 generator, V = E / 2),
                seed 5

File formats (the reference programs' own):
  MoCHy CSV  one hyperedge per line, comma-separated 0-based node ids (MoCHy-master/README.md;
             read by MoCHy-master/read_data.cpp:8-45)
  THyMe TSV  "n1,n2,...<TAB>timestamp" per line (THyMe-main/code/read_data.cpp:22-58)
  ESCHER     whitespace-separated 1-based node ids per line, '#' comments and empty lines
             skipped, a non-positive id is an error (dynamicHyperGraph/utils/utils.cpp:66-130)
"""
from __future__ import annotations

import math

import numpy as np

from . import GenParams, generate_static, generate_timestamps, pick

STATIC_C1 = GenParams(num_nodes=2_000, num_edges=5_000, alpha=2.5, kmin=2, kmax=25, zipf_s=1.2, seed=1)
STATIC_C2 = GenParams(num_nodes=1_000_000, num_edges=2_000_000, alpha=2.2, kmin=2, kmax=50,
                      zipf_s=1.1, seed=2)
DYNAMIC_C3 = GenParams(num_nodes=200_000, num_edges=500_000, alpha=2.2, kmin=2, kmax=30, zipf_s=1.1,
                       seed=3)
DYNAMIC_C3_BATCHES = 100
DYNAMIC_C3_INSERTS = 7_000
DYNAMIC_C3_DELETES = 3_000
TEMPORAL_C4 = GenParams(num_nodes=100_000, num_edges=200_000, alpha=2.2, kmin=2, kmax=20, zipf_s=1.2,
                        seed=4)
TEMPORAL_C4_DELTA_US = 3_600_000_000
SERVING_C5_GRAPHS = 64


def scaled(p: GenParams, num_edges: int, num_nodes: int | None = None, seed: int | None = None) -> GenParams:
    """Same generator family with another size (reduced instances for oracle-sized checks)."""
    return GenParams(num_nodes=num_nodes if num_nodes is not None else p.num_nodes,
                     num_edges=num_edges, alpha=p.alpha, kmin=p.kmin,
                     kmax=min(p.kmax, num_nodes if num_nodes is not None else p.num_nodes),
                     zipf_s=p.zipf_s, seed=p.seed if seed is None else seed)


def temporal_c4(num_edges: int | None = None, num_nodes: int | None = None):
    """(offsets, nodes, times) of the temporal workload (optionally reduced)."""
    p = TEMPORAL_C4 if num_edges is None else scaled(TEMPORAL_C4, num_edges, num_nodes)
    off, nodes = generate_static(p)
    times = generate_timestamps(p.num_edges, seed=p.seed)
    return off, nodes, times


def dynamic_batch(live_ids: np.ndarray, b: int, params: GenParams = DYNAMIC_C3,
                  inserts: int = DYNAMIC_C3_INSERTS, deletes: int = DYNAMIC_C3_DELETES):
    """Batch b of the dynamic workload: (deleted ids, inserted offsets, inserted nodes)."""
    seed = 3000 + b
    k = min(deletes, live_ids.size)
    dels = live_ids[pick(live_ids.size, k, seed)] if k else np.zeros(0, dtype=np.int64)
    ins = GenParams(num_nodes=params.num_nodes, num_edges=inserts, alpha=params.alpha,
                    kmin=params.kmin, kmax=params.kmax, zipf_s=params.zipf_s, seed=seed)
    off, nodes = generate_static(ins)
    return np.sort(dels.astype(np.int32)), off, nodes


def serving_c5_sizes(n_graphs: int = SERVING_C5_GRAPHS, lo: int = 10_000, hi: int = 1_000_000,
                     seed: int = 5) -> list[int]:
    """Hyperedge counts log-uniform on [lo, hi] (deterministic in seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u = rng.random(n_graphs)
    return [int(round(math.exp(math.log(lo) + x * (math.log(hi) - math.log(lo))))) for x in u]


# Reduced serving workload: the static_c2 generator family at sizes the exact engine counts per
# request in milliseconds (the full [10k, 1M] range holds graphs with > 1e13 triples each).
SERVING_REDUCED_LO, SERVING_REDUCED_HI = 1_000, 6_000


def _serving_graph(e: int, i: int, seed: int):
    """Graph i of a serving set: the static_c2 family with E = e, V = E / 2, seed * 1000 + i."""
    p = GenParams(num_nodes=max(e // 2, STATIC_C2.kmax), num_edges=e, alpha=STATIC_C2.alpha,
                  kmin=STATIC_C2.kmin, kmax=STATIC_C2.kmax, zipf_s=STATIC_C2.zipf_s,
                  seed=seed * 1000 + i)
    return generate_static(p)


def serving_graphs(n_graphs: int = SERVING_C5_GRAPHS, lo: int = SERVING_REDUCED_LO,
                   hi: int = SERVING_REDUCED_HI, seed: int = 5):
    """n_graphs hypergraphs of the static_c2 family, E_i log-uniform on [lo, hi], V_i = E_i / 2."""
    return [_serving_graph(e, i, seed) for i, e in enumerate(serving_c5_sizes(n_graphs, lo, hi, seed))]


# Largest hyperedge count of the full serving configuration's graphs served at their own size, and
# the size from which a served graph is counted and updated hub-aware (on this family the hub-aware
# count and delta are faster from ~10k hyperedges up: 30k hyperedges, count 0.085 s against 1.41 s
# by triangles, a 2% update 0.11 s against 0.36 s).
SERVING_BAND_MAX = 130_000
SERVING_HUB_MIN_EDGES = 10_000


def serving_c5_band(max_edges: int = SERVING_BAND_MAX, n_graphs: int = SERVING_C5_GRAPHS,
                    seed: int = 5):
    """The graphs of the full serving configuration (64 graphs, E log-uniform on [10k, 1M], seed 5)
    that hold at most max_edges hyperedges, each at its own size and seed (38 graphs of 10,055 to
    129,515 hyperedges for the defaults): the part of that configuration exact per-request counting
    reaches with the hub-aware count (the 129,515-hyperedge graph holds ~6.7e13 triples: 1.5 s per
    count, ~1 s per update of 2% of its hyperedges)."""
    sizes = serving_c5_sizes(n_graphs, 10_000, 1_000_000, seed)
    return [_serving_graph(e, i, seed) for i, e in enumerate(sizes) if e <= max_edges]


def serving_requests(graphs, n_requests: int, seed: int = 5000, count_frac: float = 0.5,
                     insert_frac: float = 0.7, min_batch_frac: float = 0.005,
                     max_batch_frac: float = 0.05, rate_per_s: float = 0.0):
    """A seeded request stream over `graphs`: count queries and update batches (batch sizes
    log-uniform on [min, max] x E, insert_frac of the changes are fresh generator hyperedges, the
    rest deletes of uniformly chosen live hyperedges), graphs chosen uniformly, Poisson arrival
    times at rate_per_s (all at 0 when rate_per_s is 0). Live sets are tracked as the engine
    assigns ids (inserted hyperedges get the next ids), so both a plain and a served run apply
    exactly the same updates."""
    rng = np.random.Generator(np.random.PCG64(seed))
    live = [list(range(off.size - 1)) for off, _ in graphs]
    nxt = [off.size - 1 for off, _ in graphs]
    reqs, t = [], 0.0
    for r in range(n_requests):
        if rate_per_s > 0:
            t += float(rng.exponential(1.0 / rate_per_s))
        gi = int(rng.integers(len(graphs)))
        if rng.random() < count_frac:
            reqs.append({"t": t, "graph": gi, "kind": "count"})
            continue
        e0 = graphs[gi][0].size - 1
        lo, hi = max(1.0, min_batch_frac * e0), max(1.0, max_batch_frac * e0)
        size = int(round(math.exp(math.log(lo) + rng.random() * (math.log(hi) - math.log(lo)))))
        n_ins = int(round(size * insert_frac))
        n_del = min(size - n_ins, len(live[gi]))
        pick_idx = rng.choice(len(live[gi]), size=n_del, replace=False) if n_del else []
        dels = sorted(live[gi][k] for k in pick_idx)
        keep = set(dels)
        live[gi] = [x for x in live[gi] if x not in keep]
        p = GenParams(num_nodes=max(e0 // 2, STATIC_C2.kmax), num_edges=n_ins, alpha=STATIC_C2.alpha,
                      kmin=STATIC_C2.kmin, kmax=STATIC_C2.kmax, zipf_s=STATIC_C2.zipf_s,
                      seed=seed * 100_000 + r)
        ioff, inodes = generate_static(p)
        live[gi].extend(range(nxt[gi], nxt[gi] + n_ins))
        nxt[gi] += n_ins
        reqs.append({"t": t, "graph": gi, "kind": "update", "dels": np.asarray(dels, dtype=np.int32),
                     "ins_off": ioff, "ins_nodes": inodes})
    return reqs


# ------------------------------------------------------------------------------------------
# Reference file formats
# ------------------------------------------------------------------------------------------
def write_mochy_csv(path: str, offsets: np.ndarray, nodes: np.ndarray) -> None:
    with open(path, "w") as f:
        for e in range(offsets.size - 1):
            f.write(",".join(map(str, nodes[offsets[e]:offsets[e + 1]].tolist())))
            f.write("\n")


def read_mochy_csv(path: str) -> tuple[np.ndarray, np.ndarray]:
    """MoCHy CSV; a node repeated inside a line is kept once (MoCHy-master/read_data.cpp:34)."""
    offs = [0]
    flat: list[int] = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            seen: dict[int, None] = {}
            for tok in line.split(","):
                seen.setdefault(int(tok), None)
            flat.extend(seen)
            offs.append(len(flat))
    return np.asarray(offs, dtype=np.int64), np.asarray(flat, dtype=np.int32)


def write_thyme_tsv(path: str, offsets: np.ndarray, nodes: np.ndarray, times: np.ndarray) -> None:
    with open(path, "w") as f:
        for e in range(offsets.size - 1):
            f.write(",".join(map(str, nodes[offsets[e]:offsets[e + 1]].tolist())))
            f.write("\t%d\n" % int(times[e]))


def read_thyme_tsv(path: str) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    offs = [0]
    flat: list[int] = []
    times: list[int] = []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if not line:
                continue
            ids, t = line.split("\t")[:2]
            seen: dict[int, None] = {}
            for tok in ids.split(","):
                seen.setdefault(int(tok), None)
            flat.extend(seen)
            offs.append(len(flat))
            times.append(int(float(t)))
    return (np.asarray(offs, dtype=np.int64), np.asarray(flat, dtype=np.int32),
            np.asarray(times, dtype=np.int64))


def write_escher(path: str, offsets: np.ndarray, nodes: np.ndarray) -> None:
    """ESCHER's input format: 1-based node ids, space-separated (saveHypergraphToFile)."""
    with open(path, "w") as f:
        for e in range(offsets.size - 1):
            f.write(" ".join(str(int(v) + 1) for v in nodes[offsets[e]:offsets[e + 1]]))
            f.write("\n")


def read_escher(path: str) -> tuple[np.ndarray, np.ndarray]:
    """ESCHER's input format (loadHypergraphFromFile) as 0-based node ids; repeated ids inside a
    line are kept once (the counting semantics treat hyperedges as sets)."""
    offs = [0]
    flat: list[int] = []
    with open(path) as f:
        for lineno, line in enumerate(f, 1):
            if not line.strip() or line.startswith("#"):
                continue
            seen: dict[int, None] = {}
            for tok in line.split():
                v = int(tok)
                if v <= 0:
                    raise ValueError(f"{path}:{lineno}: non-positive vertex id {v}")
                seen.setdefault(v - 1, None)
            flat.extend(seen)
            offs.append(len(flat))
    if len(offs) == 1:
        raise ValueError(f"{path}: no hyperedges")
    return np.asarray(offs, dtype=np.int64), np.asarray(flat, dtype=np.int32)
