"""GPU parity of the three-snapshot strict-increasing count (hm_count_strict_inc, the reference's
temporal-window interface) against the oracle, its brute-force definition and the reference
driver's synthetic window. Integer counts: bit-exact."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import csr, random_hypergraph

import ipdps26_amd as hm

pytestmark = pytest.mark.gpu


def snapshots(rng, num_edges, num_nodes, kmax, edits):
    off, nodes = random_hypergraph(rng, num_edges=num_edges, num_nodes=num_nodes, kmax=kmax)
    rows = [list(nodes[off[e]:off[e + 1]]) for e in range(off.size - 1)]
    layers = []
    for _ in range(3):
        snap = [list(r) for r in rows]
        for e in rng.choice(len(snap), size=min(edits, len(snap)), replace=False):
            k = int(rng.integers(1, kmax + 1))
            snap[e] = [] if rng.random() < 0.3 else list(rng.choice(num_nodes, size=min(k, num_nodes), replace=False))
        snap += [list(rng.choice(num_nodes, size=min(3, num_nodes), replace=False))
                 for _ in range(int(rng.integers(0, 20)))]
        layers.append(csr(snap))
    return layers


def test_reference_synthetic_window(gpu_ready):
    L = csr([[1, 2], [2, 3], [1, 3]])
    L2 = csr([[1, 2], [2, 3], [1, 2, 3]])
    old = hm.count_strict_inc(L, L, L).astype(np.int64)
    new = hm.count_strict_inc(L, L, L2).astype(np.int64)
    assert old[26] == 1 and old.sum() == 1
    d = new - old
    assert d[10] == 1 and d[26] == -1 and np.abs(d).sum() == 2


@pytest.mark.parametrize("seed", range(10))
def test_random_snapshots_match_oracle(gpu_ready, seed):
    rng = np.random.default_rng(9000 + seed)
    kmax = [5, 12, 40][seed % 3]  # 40: 64-bit masks
    layers = snapshots(rng, int(rng.integers(20, 1200)), int(rng.integers(kmax, 300)), kmax,
                       int(rng.integers(0, 60)))
    print(f"seed={seed} rows={[l[0].size - 1 for l in layers]}", flush=True)
    got = hm.count_strict_inc(*layers).astype(np.int64)
    assert np.array_equal(got, O.strict_inc(*layers))


@pytest.mark.parametrize("seed", range(3))
def test_tiny_snapshots_brute_force(gpu_ready, seed):
    layers = snapshots(np.random.default_rng(9100 + seed), 30, 12, 5, 10)
    assert np.array_equal(hm.count_strict_inc(*layers).astype(np.int64), O.strict_inc_brute(*layers))


def test_identical_snapshots_equal_ordered_closed_count(gpu_ready):
    # with three identical snapshots every closed triple {a, b, c} of the static graph is counted
    # exactly once (as a < b < c), so the strict-inc bins equal the closed static bins
    rng = np.random.default_rng(77)
    off, nodes = random_hypergraph(rng, num_edges=600, num_nodes=150, kmax=8)
    s = hm.count_strict_inc((off, nodes), (off, nodes), (off, nodes)).astype(np.int64)
    with hm.Graph(off, nodes) as g:
        st = g.count_static().astype(np.int64)
    closed = st.copy()
    closed[20:26] = 0
    assert np.array_equal(s, closed)


def test_empty_and_bad_inputs(gpu_ready):
    E = (np.zeros(1, dtype=np.int64), np.zeros(0, dtype=np.int32))
    assert hm.count_strict_inc(E, E, E).sum() == 0
    L = csr([[1, 2], [2, 3]])
    assert hm.count_strict_inc(L, E, L).sum() == 0
    with pytest.raises(hm.HMotifError):
        hm.count_strict_inc(csr([list(range(70))]), L, L)
