"""Batched static counts (hm_batch_count_static through the C ABI): every graph's 30 bins equal
its one-graph count (hm_count_static) and, where it finishes, the MoCHy oracle; every triangle
form (waves per centre, LDS histogram copies, persistent work queue or one workgroup per centre)
gives the same integers in every centre order; the autotuner runs once per batch shape. Integer counts: bit-exact."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import csr, random_hypergraph

import ipdps26_amd as hm
from ipdps26_amd import workloads as W

pytestmark = pytest.mark.gpu

FORMS = [(w, c, p, o) for w in (4, 8) for c in (8, 16) for p in (0, 1) for o in (0, 1, 2)]


def one_by_one(graphs):
    return np.stack([g.count_static() for g in graphs]) if graphs else np.zeros((0, 30), np.uint64)


def mixed_graphs(seed, n):
    """Small random graphs of both mask widths (max size <= 32 and > 32), with duplicates."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        kmax = [3, 8, 20, 40, 64][k % 5]
        out.append(random_hypergraph(rng, num_edges=int(rng.integers(3, 400)),
                                     num_nodes=int(rng.integers(kmax, 150)), kmax=kmax, dup_frac=0.2))
    return out


@pytest.mark.parametrize("seed", range(4))
def test_batch_matches_oracle_and_single(gpu_ready, seed):
    specs = mixed_graphs(500 + seed, 12)
    graphs = [hm.Graph(o, n) for o, n in specs]
    try:
        with hm.Batch() as b:
            got = b.count_static(graphs)
            st = b.stats()
        assert st["graphs_batched"] == len(graphs) and st["graphs_alone"] == 0
        assert np.array_equal(got, one_by_one(graphs))
        for (o, n), row in list(zip(specs, got))[:4]:
            assert np.array_equal(row.astype(np.int64), O.static_mochy(o, n))
    finally:
        for g in graphs:
            g.close()


@pytest.mark.parametrize("form", FORMS, ids=[f"w{w}c{c}p{p}o{o}" for w, c, p, o in FORMS])
def test_every_form_gives_the_same_counts(gpu_ready, form):
    specs = W.serving_graphs(16, lo=300, hi=3000, seed=21) + mixed_graphs(600, 10)
    graphs = [hm.Graph(o, n) for o, n in specs]
    try:
        expect = one_by_one(graphs)
        with hm.Batch() as b:
            b.set_form(*form)
            got = b.count_static(graphs)
            st = b.stats()
        assert (st["waves"], st["copies"], st["persistent"], st["order"], st["tuned"]) == (*form, 0)
        assert np.array_equal(got, expect)
    finally:
        for g in graphs:
            g.close()


def test_autotuner_runs_once_per_shape(gpu_ready):
    graphs = [hm.Graph(o, n) for o, n in W.serving_graphs(24, lo=500, hi=2500, seed=22)]
    try:
        expect = one_by_one(graphs)
        with hm.Batch() as b:
            first = b.count_static(graphs)
            s1 = b.stats()
            second = b.count_static(graphs)
            s2 = b.stats()
        assert s1["tuned"] == 1 and s2["tuned"] == 0
        keys = ("waves", "copies", "persistent", "order")
        assert [s1[k] for k in keys] == [s2[k] for k in keys]
        assert np.array_equal(first, expect) and np.array_equal(second, expect)
    finally:
        for g in graphs:
            g.close()


def test_small_large_and_updated_graphs(gpu_ready):
    """Graphs with fewer than 3 hyperedges, one too large for one LDS table (counted alone by
    the tiled path), the same graph twice, and a graph after deletes and inserts."""
    tiny = [csr([]), csr([[1, 2]]), csr([[1, 2], [2, 3]])]
    # 17,000 hyperedges of up to 40 nodes: more than one 64-bit mask table holds
    big = hm.generate_static(hm.GenParams(num_nodes=100_000, num_edges=17_000, alpha=2.2, kmin=2,
                                          kmax=40, zipf_s=0.8, seed=7))
    off, nodes = W.serving_graphs(1, lo=900, hi=900, seed=23)[0]
    upd = hm.Graph(off, nodes)
    rng = np.random.default_rng(23)
    o2, n2 = csr([sorted(rng.choice(400, size=5, replace=False).tolist()) for _ in range(30)])
    upd.update(sorted(rng.choice(off.size - 1, size=60, replace=False).tolist()), o2, n2)
    graphs = [hm.Graph(*t) for t in tiny] + [hm.Graph(*big), upd]
    mid = hm.Graph(*W.serving_graphs(1, lo=2000, hi=2000, seed=24)[0])
    graphs += [mid, mid]
    try:
        expect = one_by_one(graphs)
        assert not expect[:3].any()
        with hm.Batch() as b:
            got = b.count_static(graphs)
            st = b.stats()
        assert st["graphs_alone"] == 1 and st["graphs_batched"] == 3
        assert np.array_equal(got, expect)
    finally:
        for g in graphs[:-1]:
            g.close()


def test_empty_batch(gpu_ready):
    with hm.Batch() as b:
        assert b.count_static([]).shape == (0, 30)
