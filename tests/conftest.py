import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(__file__))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


def load_golden(kind: str):
    path = os.path.join(GOLDEN, f"{kind}_golden.json")
    if not os.path.exists(path):
        return []
    with open(path) as f:
        return json.load(f)


def random_hypergraph(rng: np.random.Generator, num_edges: int, num_nodes: int, kmax: int,
                      dup_frac: float = 0.15):
    """Small random hypergraph with skewed node popularity, duplicate and nested hyperedges."""
    edges = []
    pop = 1.0 / np.arange(1, num_nodes + 1) ** 0.9
    pop /= pop.sum()
    for _ in range(num_edges):
        r = rng.random()
        if edges and r < dup_frac:
            edges.append(list(edges[rng.integers(len(edges))]))  # duplicate node set
        elif edges and r < 2 * dup_frac:
            base = edges[rng.integers(len(edges))]
            k = max(1, int(rng.integers(1, len(base) + 1)))
            edges.append(list(rng.choice(base, size=k, replace=False)))  # subset
        else:
            k = int(rng.integers(1, kmax + 1))
            k = min(k, num_nodes)
            edges.append(list(rng.choice(num_nodes, size=k, replace=False, p=pop)))
    offsets = np.zeros(len(edges) + 1, dtype=np.int64)
    flat = []
    for i, e in enumerate(edges):
        flat.extend(int(x) for x in e)
        offsets[i + 1] = len(flat)
    return offsets, np.asarray(flat, dtype=np.int32)


def csr(edges):
    offsets = np.zeros(len(edges) + 1, dtype=np.int64)
    flat = []
    for i, e in enumerate(edges):
        flat.extend(int(x) for x in e)
        offsets[i + 1] = len(flat)
    return offsets, np.asarray(flat, dtype=np.int32)


@pytest.fixture(scope="session")
def gpu_ready():
    """Initialises libhmotif on device 0 (gfx950); tests marked gpu use it."""
    import ipdps26_amd as hm

    hm.init(0)
    return hm
