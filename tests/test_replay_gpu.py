"""hipGraph replay of the static count (hm_graph_set_replay through the C ABI): the captured
count of an unchanged graph, replayed many times with other eager work on the device between
replays, gives the eager count every time; updates and tile changes drop the capture. Integer
counts: bit-exact."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import csr, random_hypergraph

import ipdps26_amd as hm
from ipdps26_amd import workloads as W

pytestmark = pytest.mark.gpu


def graphs_under_test():
    rng = np.random.default_rng(808)
    narrow = random_hypergraph(rng, num_edges=400, num_nodes=120, kmax=12, dup_frac=0.2)
    wide = random_hypergraph(rng, num_edges=300, num_nodes=200, kmax=60, dup_frac=0.1)
    mid = W.serving_graphs(1, lo=3000, hi=3000, seed=81)[0]
    return {"narrow": narrow, "wide": wide, "mid": mid}


@pytest.mark.parametrize("name", ["narrow", "wide", "mid"])
@pytest.mark.parametrize("tile", [0, 64])
def test_replays_equal_eager(gpu_ready, name, tile):
    off, nodes = graphs_under_test()[name]
    other = hm.Graph(*W.serving_graphs(1, lo=1500, hi=1500, seed=82)[0])
    with hm.Graph(off, nodes) as g, hm.Graph(off, nodes) as ref:
        g.set_tile(tile)
        ref.set_tile(tile)
        expect = ref.count_static()
        st_expect = ref.stats()
        g.set_replay(True)
        first = g.count_static()
        assert g.replay_state() == (True, False)  # eager count, then captured
        for k in range(6):
            other.count_static()  # eager work of another graph between replays
            got = g.count_static()
            assert g.replay_state() == (True, True), k
            assert np.array_equal(got, expect), f"replay {k}"
        assert np.array_equal(first, expect)
        assert g.stats() == st_expect
    other.close()
    if name == "narrow" and tile == 0:
        assert np.array_equal(expect.astype(np.int64), O.static_mochy(off, nodes))


def test_update_and_tile_change_drop_the_capture(gpu_ready):
    off, nodes = W.serving_graphs(1, lo=2000, hi=2000, seed=83)[0]
    rng = np.random.default_rng(83)
    with hm.Graph(off, nodes) as g:
        g.set_replay(True)
        g.count_static()
        g.count_static()
        assert g.replay_state() == (True, True)
        ins = csr([sorted(rng.choice(1000, size=4, replace=False).tolist()) for _ in range(20)])
        g.update(sorted(rng.choice(off.size - 1, size=30, replace=False).tolist()), *ins)
        assert g.replay_state()[0] is False
        after = g.count_static()
        assert g.replay_state() == (True, False)
        assert np.array_equal(g.count_static(), after) and g.replay_state() == (True, True)
        g.set_tile(128)
        assert g.replay_state()[0] is False
        assert np.array_equal(g.count_static(), after)
        assert np.array_equal(g.count_static(), after) and g.replay_state() == (True, True)
        g.set_replay(False)
        assert np.array_equal(g.count_static(), after) and g.replay_state() == (False, False)
