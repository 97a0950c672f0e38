"""Captured updates (hm_graph_set_update_capture through the C ABI): an update with a delta run as
one hipGraph gives, batch after batch, the same delta, removed / added triples, scanned row
entries, assigned ids and store as the eager update of the same graph — including batches that
outgrow the plan (captured again), rows that outgrow it on the device (addition phase rerun
eagerly) and a mask-width change. Integer counts: bit-exact."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import csr, random_hypergraph

import ipdps26_amd as hm
from ipdps26_amd import workloads as W

pytestmark = pytest.mark.gpu


def rows_of(off, nodes):
    return [list(nodes[off[e]:off[e + 1]]) for e in range(off.size - 1)]


def random_batch(rng, rows, live, n_del, n_ins, num_nodes, kmax):
    dels = sorted(rng.choice(sorted(live), size=min(n_del, len(live)), replace=False).tolist())
    new_rows = []
    for _ in range(n_ins):
        if rng.random() < 0.2 and live:  # duplicate of a live hyperedge
            new_rows.append(list(rows[int(rng.choice(sorted(live)))]))
        else:
            k = int(rng.integers(1, kmax + 1))
            new_rows.append(list(rng.choice(num_nodes, size=min(k, num_nodes), replace=False)))
    return dels, new_rows


def same_update(ge, gc, dels, ioff, inodes, tag):
    """Applies one batch to the eager graph ge and the captured graph gc; checks they agree."""
    ids_e, de = ge.update(dels, ioff, inodes)
    ids_c, dc = gc.update(dels, ioff, inodes)
    assert np.array_equal(ids_e, ids_c), tag
    assert np.array_equal(de, dc), tag
    se, sc = ge.update_stats(), gc.update_stats()
    for k in ("removed", "added", "row_entries"):
        assert se[k] == sc[k], (tag, k)
    assert se["captured"] == 0
    return dc, sc


@pytest.mark.parametrize("seed", range(6))
def test_captured_batches_equal_eager(gpu_ready, seed):
    rng = np.random.default_rng(7100 + seed)
    nv, kmax = int(rng.integers(15, 90)), [4, 9, 30][seed % 3]
    off, nodes = random_hypergraph(rng, num_edges=int(rng.integers(30, 300)), num_nodes=nv, kmax=kmax)
    rows = rows_of(off, nodes)
    live = set(range(len(rows)))
    with hm.Graph(off, nodes) as ge, hm.Graph(off, nodes) as gc:
        gc.set_update_capture(True)
        before = O.static_mochy(*csr(rows))
        for b in range(6):
            n_del = int(rng.integers(0, 20)) if b != 2 else 0
            n_ins = int(rng.integers(0, 20)) if b != 3 else 0
            dels, new_rows = random_batch(rng, rows, live, n_del, n_ins, nv, kmax)
            ioff, inodes = csr(new_rows)
            print(f"seed={seed} batch={b} E={len(rows)} del={len(dels)} ins={len(new_rows)}", flush=True)
            delta, st = same_update(ge, gc, dels, ioff, inodes, (seed, b))
            assert st["captured"] == 1, (seed, b)
            ids = list(range(len(rows), len(rows) + len(new_rows)))
            for x in dels:
                rows[x] = []
                live.discard(x)
            rows.extend(new_rows)
            live.update(ids)
            after = O.static_mochy(*csr(rows))
            assert np.array_equal(delta, after - before), (seed, b)
            before = after
        state = gc.update_capture_state()
        assert state["on"] == 1 and state["disabled"] == 0 and state["captured"] == 1
        assert state["last_captured"] == 1
        # the captured store counts like the eager one and like the oracle
        assert np.array_equal(gc.count_static(), ge.count_static())
        assert np.array_equal(gc.count_static().astype(np.int64), before)
        assert gc.stats() == ge.stats()


def test_plan_grows_when_batches_outgrow_it(gpu_ready):
    """Batches larger than the first, and ids beyond the first plan, capture again."""
    off, nodes = W.serving_graphs(1, lo=1200, hi=1200, seed=91)[0]
    rng = np.random.default_rng(91)
    nv = int(nodes.max()) + 1
    rows = rows_of(off, nodes)
    live = set(range(len(rows)))
    with hm.Graph(off, nodes) as ge, hm.Graph(off, nodes) as gc:
        gc.set_update_capture(True)
        sizes = [(5, 10), (5, 10), (40, 120), (10, 30), (200, 600), (3, 5), (0, 900)]
        for b, (n_del, n_ins) in enumerate(sizes):
            dels, new_rows = random_batch(rng, rows, live, n_del, n_ins, nv + 50, 12)
            ioff, inodes = csr(new_rows)
            print(f"batch={b} E={len(rows)} del={len(dels)} ins={len(new_rows)}", flush=True)
            same_update(ge, gc, dels, ioff, inodes, b)
            for x in dels:
                rows[x] = []
                live.discard(x)
            live.update(range(len(rows), len(rows) + len(new_rows)))
            rows.extend(new_rows)
            st = gc.update_capture_state()
            assert st["last_captured"] == 1, b
            assert st["id_capacity"] >= len(rows)
        assert gc.update_capture_state()["captures"] >= 3
        assert np.array_equal(gc.count_static(), ge.count_static())
        with hm.Graph(*csr(rows)) as fresh:
            assert np.array_equal(fresh.count_static(), gc.count_static())


def test_row_overflow_reruns_the_addition_phase(gpu_ready):
    """Inserted hyperedges that overlap the whole graph outgrow the plan's row capacity: the
    addition phase is skipped on the device and rerun eagerly; the next batch captures a larger
    plan. Deltas stay equal to the eager ones and to the oracle."""
    rng = np.random.default_rng(92)
    off, nodes = random_hypergraph(rng, num_edges=300, num_nodes=60, kmax=5)
    rows = rows_of(off, nodes)
    with hm.Graph(off, nodes) as ge, hm.Graph(off, nodes) as gc:
        gc.set_update_capture(True)
        small = [sorted(rng.choice(60, size=3, replace=False).tolist()) for _ in range(4)]
        same_update(ge, gc, [0, 1], *csr(small), "warm")
        rows[0] = []
        rows[1] = []
        rows.extend(small)
        cap = gc.update_capture_state()["row_capacity"]
        hubs = [sorted(rng.choice(60, size=40, replace=False).tolist()) for _ in range(60)]
        before = O.static_mochy(*csr(rows))
        delta, st = same_update(ge, gc, [2, 3], *csr(hubs), "hubs")
        assert st["captured"] == 1 and st["rows_overflow"] == 1
        assert gc.update_capture_state()["last_overflow"] == 1
        rows[2] = []
        rows[3] = []
        rows.extend(hubs)
        after = O.static_mochy(*csr(rows))
        assert np.array_equal(delta, after - before)
        more = [sorted(rng.choice(60, size=4, replace=False).tolist()) for _ in range(6)]
        delta, st = same_update(ge, gc, [4], *csr(more), "after")
        assert st["captured"] == 1 and st["rows_overflow"] == 0
        assert gc.update_capture_state()["row_capacity"] > cap
        rows[4] = []
        rows.extend(more)
        assert np.array_equal(delta, O.static_mochy(*csr(rows)) - after)
        assert np.array_equal(gc.count_static(), ge.count_static())


def test_mask_width_change_and_switching_off(gpu_ready):
    """A batch that brings the first hyperedge of more than 32 nodes captures a 64-bit plan;
    switching capture off runs the next update eagerly."""
    rng = np.random.default_rng(93)
    off, nodes = random_hypergraph(rng, num_edges=200, num_nodes=100, kmax=20)
    rows = rows_of(off, nodes)
    with hm.Graph(off, nodes) as ge, hm.Graph(off, nodes) as gc:
        gc.set_update_capture(True)
        b1 = [sorted(rng.choice(100, size=8, replace=False).tolist()) for _ in range(5)]
        same_update(ge, gc, [5, 6], *csr(b1), "narrow")
        n1 = gc.update_capture_state()["captures"]
        b2 = [sorted(rng.choice(100, size=45, replace=False).tolist()) for _ in range(3)]
        _, st = same_update(ge, gc, [7], *csr(b2), "wide")
        assert st["captured"] == 1
        assert gc.update_capture_state()["captures"] == n1 + 1
        b3 = [sorted(rng.choice(100, size=6, replace=False).tolist()) for _ in range(4)]
        same_update(ge, gc, [8, 9, 10], *csr(b3), "wide again")
        gc.set_update_capture(False)
        _, st = same_update(ge, gc, [11], *csr(b1[:2]), "off")
        assert st["captured"] == 0
        assert gc.update_capture_state()["captured"] == 0
        for x in (5, 6, 7, 8, 9, 10, 11):
            rows[x] = []
        rows.extend(b1 + b2 + b3 + b1[:2])
        assert np.array_equal(gc.count_static().astype(np.int64), O.static_mochy(*csr(rows)))


def test_c3_family_chain_captured(gpu_ready):
    """Reduced dynamic workload (c3 generator at 4,000 hyperedges, 6 batches of 140 inserts and
    60 deletes): captured deltas chain to the recount, and count replay stays correct beside it."""
    p = W.scaled(W.DYNAMIC_C3, 4000, 1600)
    off, nodes = hm.generate_static(p)
    with hm.Graph(off, nodes) as g:
        g.set_update_capture(True)
        g.set_replay(True)
        live = np.arange(off.size - 1, dtype=np.int64)
        counts = g.count_static().astype(np.int64)
        for b in range(1, 7):
            dels, ioff, inodes = W.dynamic_batch(live, b, p, inserts=140, deletes=60)
            ids, delta = g.update(dels, ioff, inodes)
            assert g.update_stats()["captured"] == 1, b
            after = g.count_static().astype(np.int64)
            assert np.array_equal(after, g.count_static().astype(np.int64))  # replayed count
            assert np.array_equal(counts + delta, after), f"batch {b}"
            counts = after
            live = np.concatenate([np.setdiff1d(live, dels.astype(np.int64)), ids.astype(np.int64)])
