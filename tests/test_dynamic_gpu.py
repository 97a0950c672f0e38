"""GPU parity of the dynamic path (hm_graph_update, hm_count_delta through the C ABI): every
batch's incremental delta equals the difference of two full recounts, by the oracle and by the
engine itself, and the updated store equals a graph rebuilt from scratch. Integer counts: exact."""
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
        r = rng.random()
        if r < 0.2 and live:  # duplicate of a live hyperedge
            new_rows.append(list(rows[int(rng.choice(sorted(live)))]))
        else:
            k = int(rng.integers(1, kmax + 1))
            new_rows.append(list(rng.choice(num_nodes, size=min(k, num_nodes), replace=False)))
    return dels, new_rows


@pytest.mark.parametrize("seed", range(8))
def test_batches_match_oracle_recount(gpu_ready, seed):
    rng = np.random.default_rng(5000 + seed)
    nv, kmax = int(rng.integers(10, 80)), [4, 8, 40][seed % 3]
    off, nodes = random_hypergraph(rng, num_edges=int(rng.integers(20, 250)), num_nodes=nv, kmax=kmax)
    rows = rows_of(off, nodes)
    live = set(range(len(rows)))
    g = hm.Graph(off, nodes)
    before = O.static_mochy(*csr(rows))
    for b in range(4):
        dels, new_rows = random_batch(rng, rows, live, int(rng.integers(0, 15)),
                                      int(rng.integers(0, 15)), nv, kmax)
        ioff, inodes = csr(new_rows)
        ids, delta = g.update(dels, ioff, inodes)
        assert ids.tolist() == list(range(len(rows), len(rows) + len(new_rows)))
        for x in dels:
            rows[x] = []
            live.discard(x)
        mid = int(O.static_mochy(*csr(rows)).sum())  # deleted rows gone, nothing inserted yet
        rows.extend(new_rows)
        live.update(ids.tolist())
        after = O.static_mochy(*csr(rows))
        print(f"seed={seed} batch={b} E={len(rows)}", flush=True)
        assert np.array_equal(delta, after - before)
        us = g.update_stats()
        assert (us["removed"], us["added"]) == (int(before.sum()) - mid, int(after.sum()) - mid)
        assert np.array_equal(g.count_static().astype(np.int64), after)
        before = after
    g.close()


def test_snapshot_delta_with_id_reuse(gpu_ready):
    """The reference's batch reuses deleted ids for inserts (dynamicHyperGraph/src/main.cu:196-199)."""
    rng = np.random.default_rng(77)
    off, nodes = random_hypergraph(rng, num_edges=180, num_nodes=40, kmax=7)
    old_rows = rows_of(off, nodes)
    dels = sorted(rng.choice(len(old_rows), size=25, replace=False).tolist())
    new_rows = [list(r) for r in old_rows]
    ins = dels[:15] + [len(old_rows) + k for k in range(10)]
    for x in dels:
        new_rows[x] = []
    new_rows.extend([[] for _ in range(10)])
    for x in ins:
        k = int(rng.integers(1, 7))
        new_rows[x] = list(rng.choice(40, size=k, replace=False))
    old_rows.extend([[] for _ in range(10)])  # same id space in both snapshots
    with hm.Graph(*csr(old_rows)) as go, hm.Graph(*csr(new_rows)) as gn:
        delta = hm.count_delta(go, gn, dels, ins)
    assert np.array_equal(delta, O.static_mochy(*csr(new_rows)) - O.static_mochy(*csr(old_rows)))


def test_c3_family_chain_against_recount(gpu_ready):
    """Reduced dynamic workload: c3 generator at 4,000 hyperedges, 6 batches of 200 changes
    (70% inserts of fresh generator hyperedges, 30% deletes of uniformly chosen live ones)."""
    p = W.scaled(W.DYNAMIC_C3, 4000, 1600)
    off, nodes = hm.generate_static(p)
    g = hm.Graph(off, nodes)
    live = np.arange(off.size - 1, dtype=np.int64)
    counts = g.count_static().astype(np.int64)
    assert np.array_equal(counts, O.static_mochy(off, nodes))
    rows = rows_of(off, nodes)
    for b in range(1, 7):
        dels, ioff, inodes = W.dynamic_batch(live, b, p, inserts=140, deletes=60)
        ids, delta = g.update(dels, ioff, inodes)
        counts_after = g.count_static().astype(np.int64)
        assert np.array_equal(counts + delta, counts_after), f"batch {b}"
        counts = counts_after
        live = np.setdiff1d(live, dels.astype(np.int64))
        live = np.concatenate([live, ids.astype(np.int64)])
        for x in dels:
            rows[x] = []
        rows.extend(rows_of(ioff, inodes))
    assert np.array_equal(counts, O.static_mochy(*csr(rows)))
    # the updated store equals a store built from scratch
    with hm.Graph(*csr(rows)) as fresh:
        assert np.array_equal(fresh.count_static().astype(np.int64), counts)
    g.close()


def test_update_rejects_bad_ids(gpu_ready):
    off, nodes = csr([[0, 1], [1, 2], [2, 3]])
    with hm.Graph(off, nodes) as g:
        with pytest.raises(hm.HMotifError):
            g.update([5])
        g.update([1])
        with pytest.raises(hm.HMotifError):
            g.update([1])  # already deleted
        ids, delta = g.update([], *csr([[3, 4]]))
        assert ids.tolist() == [3]


@pytest.mark.parametrize("seed", range(6))
def test_tiled_delta_matches_oracle(gpu_ready, seed):
    """The tiled delta (forced tiles on small graphs): every batch against oracle recounts."""
    rng = np.random.default_rng(5600 + seed)
    nv, kmax = int(rng.integers(10, 80)), [4, 8, 40][seed % 3]
    tile = [1, 64, 97][seed % 3]
    off, nodes = random_hypergraph(rng, num_edges=int(rng.integers(20, 250)), num_nodes=nv, kmax=kmax)
    rows = rows_of(off, nodes)
    live = set(range(len(rows)))
    g = hm.Graph(off, nodes)
    g.set_tile(tile)
    before = O.static_mochy(*csr(rows))
    for b in range(3):
        dels, new_rows = random_batch(rng, rows, live, int(rng.integers(0, 15)),
                                      int(rng.integers(0, 15)), nv, kmax)
        ids, delta = g.update(dels, *csr(new_rows))
        for x in dels:
            rows[x] = []
            live.discard(x)
        mid = int(O.static_mochy(*csr(rows)).sum())
        rows.extend(new_rows)
        live.update(ids.tolist())
        after = O.static_mochy(*csr(rows))
        print(f"seed={seed} tile={tile} batch={b} E={len(rows)}", flush=True)
        assert np.array_equal(delta, after - before)
        us = g.update_stats()
        assert (us["removed"], us["added"]) == (int(before.sum()) - mid, int(after.sum()) - mid)
        before = after
    g.close()


def test_tiled_snapshot_delta(gpu_ready):
    rng = np.random.default_rng(78)
    off, nodes = random_hypergraph(rng, num_edges=300, num_nodes=50, kmax=7)
    old_rows = rows_of(off, nodes)
    dels = sorted(rng.choice(len(old_rows), size=30, replace=False).tolist())
    new_rows = [list(r) for r in old_rows]
    ins = dels[:10] + [len(old_rows) + k for k in range(12)]
    for x in dels:
        new_rows[x] = []
    new_rows.extend([[] for _ in range(12)])
    for x in ins:
        new_rows[x] = list(rng.choice(50, size=int(rng.integers(1, 7)), replace=False))
    old_rows.extend([[] for _ in range(12)])
    with hm.Graph(*csr(old_rows)) as go, hm.Graph(*csr(new_rows)) as gn:
        go.set_tile(50)
        gn.set_tile(50)
        delta = hm.count_delta(go, gn, dels, ins)
    assert np.array_equal(delta, O.static_mochy(*csr(new_rows)) - O.static_mochy(*csr(old_rows)))


def test_delta_beyond_one_table(gpu_ready):
    """40,000 hyperedge ids (more than one LDS table): the automatic path tiles; batches of the
    dynamic recipe, each delta against two recounts (the static count, itself checked against the
    oracle at this size in test_static_tiled_gpu) and the final graph against the oracle."""
    p = hm.GenParams(num_nodes=400_000, num_edges=40_000, alpha=2.8, kmin=2, kmax=12, zipf_s=0.7,
                     seed=14)
    off, nodes = hm.generate_static(p)
    g = hm.Graph(off, nodes)
    live = np.arange(off.size - 1, dtype=np.int64)
    rows = rows_of(off, nodes)
    counts = g.count_static().astype(np.int64)
    for b in range(1, 4):
        dels, ioff, inodes = W.dynamic_batch(live, b, p, inserts=560, deletes=240)
        ids, delta = g.update(dels, ioff, inodes)
        after = g.count_static().astype(np.int64)
        assert np.array_equal(counts + delta, after), f"batch {b}"
        counts = after
        live = np.concatenate([np.setdiff1d(live, dels.astype(np.int64)), ids.astype(np.int64)])
        for x in dels:
            rows[x] = []
        rows.extend(rows_of(ioff, inodes))
    assert np.array_equal(counts, O.static_mochy(*csr(rows)))
    g.close()


def expected_store(rows, num_nodes):
    """H2V (rows sorted) and V2H (hyperedge ids ascending per node) of a list of rows."""
    sizes = np.array([r.size for r in rows], dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    nodes = np.concatenate([np.sort(r) for r in rows]).astype(np.int32)
    eid = np.repeat(np.arange(len(rows), dtype=np.int32), sizes)
    order = np.argsort(nodes, kind="stable")  # edge ids stay ascending within a node
    vcnt = np.bincount(nodes, minlength=num_nodes)
    voff = np.concatenate([[0], np.cumsum(vcnt)]).astype(np.int64)
    return off, nodes, voff, eid[order]


def test_full_c3_store_updates(gpu_ready):
    """The dynamic workload at full size, store only: the 500,000-hyperedge base and its 100
    batches of 7,000 fresh generator hyperedges and 3,000 deletes of uniformly chosen live ones.
    (Its deltas are not counted here: the base's largest node sits in ~199k hyperedges.) After
    batches 1, 50 and 100 the device store, H2V and V2H, equals the rows the host keeps."""
    p = W.DYNAMIC_C3
    off, nodes = hm.generate_static(p)
    E = off.size - 1
    rows = [nodes[off[e]:off[e + 1]] for e in range(E)]
    live = np.arange(E, dtype=np.int64)
    empty = np.zeros(0, dtype=np.int32)
    ms_store = []
    with hm.Graph(off, nodes) as g:
        for b in range(1, W.DYNAMIC_C3_BATCHES + 1):
            dels, ioff, inodes = W.dynamic_batch(live, b)
            ids, _ = g.update(dels, ioff, inodes, delta=False)
            ms_store.append(g.phase_ms("store"))
            assert ids.tolist() == list(range(len(rows), len(rows) + ioff.size - 1))
            for x in dels:
                rows[int(x)] = empty
            rows.extend(inodes[ioff[k]:ioff[k + 1]] for k in range(ioff.size - 1))
            live = np.concatenate([np.setdiff1d(live, dels.astype(np.int64)), ids.astype(np.int64)])
            if b in (1, 50, W.DYNAMIC_C3_BATCHES):
                got = g.read_csr()
                want = expected_store(rows, got[2].size - 1)
                for name, x, y in zip(("offsets", "nodes", "voffsets", "vedges"), got, want):
                    assert np.array_equal(x, y), f"batch {b}: {name}"
                print(f"batch {b}: {len(rows)} ids, {int(got[0][-1])} incidences, store update "
                      f"{ms_store[-1]:.3f} ms", flush=True)
    print(f"store update per batch: mean {np.mean(ms_store):.3f} ms, max {np.max(ms_store):.3f} ms",
          flush=True)


@pytest.mark.parametrize("edges", [14_000, 24_000])
def test_delta_one_group_per_cu_equals_recounts(gpu_ready, edges):
    """Graphs whose delta table leaves one 16-wave workgroup per CU (about 11.5k to 28k ids: the
    sixteen-load form): every batch's delta equals the difference of two recounts."""
    p = W.scaled(W.DYNAMIC_C3, edges, int(edges * 0.4))
    off, nodes = hm.generate_static(p)
    live = np.arange(off.size - 1, dtype=np.int64)
    with hm.Graph(off, nodes) as g:
        before = g.count_static().astype(np.int64)
        for b in range(1, 4):
            dels, ioff, inodes = W.dynamic_batch(live, b, p, inserts=210, deletes=90)
            ids, delta = g.update(dels, ioff, inodes)
            after = g.count_static().astype(np.int64)
            print(f"E={edges} batch {b}: touched {g.update_stats()['removed'] + g.update_stats()['added']}",
                  flush=True)
            assert np.array_equal(before + delta, after), f"batch {b}"
            before = after
            live = np.concatenate([np.setdiff1d(live, dels.astype(np.int64)), ids.astype(np.int64)])
