"""GPU parity of the hub-aware static count (hm_graph_set_hubs + hm_count_static, through the C
ABI) against the oracle, brute force, the reference's golden vectors and the triangle path.
Integer counts: bit-exact."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from conftest import GOLDEN, csr, load_golden, random_hypergraph

import ipdps26_amd as hm
from ipdps26_amd import workloads as W

pytestmark = pytest.mark.gpu


def fits_one_table(off, nodes):
    """Mirror of the library's rule: mask table + 8-byte neighbour list (k_hub_centre) and mask
    table + light sizes (k_hub_tri) within 160 KB of LDS, 12 KB kept for static counters."""
    E = off.size - 1
    wide = int(np.diff(off).max()) > 32
    mask = 8 if wide else 4
    return max(E * (mask + 8), E * (mask + 1)) + 12 * 1024 <= 160 * 1024


OPEN = slice(20, 26)


def wedge_identity(c, st):
    """sum(open) + 3 sum(closed) = sum_a C(d_a, 2): every connected triple is one wedge if open
    and three if closed; the projected degrees come from the projection kernels alone."""
    closed = int(c.sum() - c[OPEN].sum())
    return int(c[OPEN].sum()) + 3 * closed == st["wedges"]


def count(off, nodes, hubs):
    with hm.Graph(off, nodes) as g:
        g.set_hubs(hubs)
        c = g.count_static().astype(np.int64)
        if g.hub_stats()["hubs"] > 0:
            assert wedge_identity(c, g.stats())
        return c, g.hub_stats()


@pytest.mark.parametrize("seed", range(12))
def test_random_matches_oracle(gpu_ready, seed):
    rng = np.random.default_rng(5000 + seed)
    kmax = [4, 8, 20, 40][seed % 4]  # 40: 64-bit light masks
    off, nodes = random_hypergraph(rng, num_edges=int(rng.integers(3, 400)),
                                   num_nodes=int(rng.integers(kmax, 120)), kmax=kmax, dup_frac=0.2)
    want = O.static_mochy(off, nodes)
    for h in (1, 2, 3, 5, 8, 16, 23):
        print(f"seed={seed} E={off.size - 1} hubs={h}", flush=True)
        c, st = count(off, nodes, h)
        assert np.array_equal(c, want), h
        assert st["hubs"] == min(h, int(nodes.max()) + 1)


@pytest.mark.parametrize("seed", range(4))
def test_tiny_matches_brute_force(gpu_ready, seed):
    rng = np.random.default_rng(6000 + seed)
    off, nodes = random_hypergraph(rng, num_edges=60, num_nodes=15, kmax=6, dup_frac=0.3)
    want = O.static_brute(off, nodes)
    for h in (1, 4, 15):
        assert np.array_equal(count(off, nodes, h)[0], want), h


@pytest.mark.parametrize("name", ["tiny_dense_e60", "c1_family_e800", "c2_family_e3000",
                                  "c3_family_e6000", "c2_family_e8000", "c5_family_e10000", "dblp"])
def test_reference_goldens(gpu_ready, name):
    g = [x for x in load_golden("static") if x["name"] == name][0]
    if g["kind"] == "fixture":
        off, nodes = W.read_mochy_csv(os.path.join(GOLDEN, g["file"]))
    else:
        off, nodes = hm.generate_static(hm.GenParams(**g["params"]))
    for h in (4, 12, -1):
        print(f"golden={name} hubs={h}", flush=True)
        c, _ = count(off, nodes, h)
        assert hm.printed_motifs(c) == g["printed_motifs"], h


def test_static_c1_all_bins(gpu_ready):
    g = json.load(open(os.path.join(GOLDEN, "static_c1_30bins.json")))
    off, nodes = hm.generate_static(W.STATIC_C1)
    for h in (2, 8, 16, -1):
        c, st = count(off, nodes, h)
        print(f"static_c1 hubs={h} stats={st}", flush=True)
        assert c.tolist() == g["counts30"], h
        assert st["hubs"] > 0


def test_equals_triangle_path_c2_family(gpu_ready):
    off, nodes = hm.generate_static(W.scaled(W.STATIC_C2, 9_000, 4_500, seed=21))
    assert fits_one_table(off, nodes)
    plain, _ = count(off, nodes, 0)
    for h in (8, 23, -1):
        print(f"c2 family 9k hubs={h}", flush=True)
        assert np.array_equal(count(off, nodes, h)[0], plain), h


def test_after_deletes_and_inserts(gpu_ready):
    rng = np.random.default_rng(77)
    off, nodes = random_hypergraph(rng, num_edges=300, num_nodes=60, kmax=10)
    with hm.Graph(off, nodes) as g:
        g.update(np.arange(0, 300, 7, dtype=np.int32), *csr([[1, 2, 3], [2, 3, 4, 50], [0, 1]]))
        plain = g.count_static().astype(np.int64)
        g.set_hubs(6)
        assert np.array_equal(g.count_static().astype(np.int64), plain)
        st = g.stats()
        assert wedge_identity(g.count_static().astype(np.int64), st)


def test_closed_forms(gpu_ready):
    # star sharing one node: C(k, 3) in bin 3; the hub is node 0
    k = 500
    c, _ = count(*csr([[0, 10_000 + i] for i in range(k)]), 1)
    assert c[3] == k * (k - 1) * (k - 2) // 6 and c.sum() == c[3]
    c, _ = count(*csr([[4, 5, 6]] * 3), 2)
    assert c[0] == 1 and c.sum() == 1
    c, _ = count(*csr([[1, 3, 10], [1, 2, 11], [2, 3, 12]]), 1)
    assert c[29] == 1 and c.sum() == 1
    assert count(*csr([[0, 1], [1, 2]]), 3)[0].sum() == 0


@pytest.mark.parametrize("seed", range(6))
def test_forced_tiles_match_oracle(gpu_ready, seed):
    rng = np.random.default_rng(7000 + seed)
    kmax = [6, 40][seed % 2]
    off, nodes = random_hypergraph(rng, num_edges=int(rng.integers(50, 600)),
                                   num_nodes=int(rng.integers(kmax, 150)), kmax=kmax, dup_frac=0.2)
    want = O.static_mochy(off, nodes)
    for tile in (64, 256, 1000):
        for h in (1, 4, 12):
            print(f"seed={seed} E={off.size - 1} tile={tile} hubs={h}", flush=True)
            with hm.Graph(off, nodes) as g:
                g.set_tile(tile)
                g.set_hubs(h)
                assert np.array_equal(g.count_static().astype(np.int64), want), (tile, h)


@pytest.mark.parametrize("seed", range(6))
def test_parts_over_centre_ranges_sum_to_the_count(gpu_ready, seed):
    """hm_graph_count_hub_part over a partition of the light ranks (the other parts once) sums to
    the oracle's bins; the light triangles of the ranges sum to the full count's; the projection
    statistics of the call with the other parts satisfy the wedge identity."""
    rng = np.random.default_rng(9100 + seed)
    kmax = [4, 12, 40][seed % 3]
    off, nodes = random_hypergraph(rng, num_edges=int(rng.integers(150, 400)),
                                   num_nodes=int(rng.integers(kmax + 10, 90)), kmax=kmax, dup_frac=0.1)
    E = off.size - 1
    hubs = [2, 5, 9][seed % 3]
    want = O.static_mochy(off, nodes)
    cuts = sorted(set([0, E] + rng.integers(0, E, 3).tolist()))
    with hm.Graph(off, nodes) as g:
        if seed % 2:
            g.set_tile(64)
        g.set_hubs(hubs)
        full = g.count_static().astype(np.int64)
        full_lt = g.hub_stats()["light_triangles"]
        total, lt = np.zeros(hm.STATIC_BINS, dtype=np.uint64), 0
        for i, (a, b) in enumerate(zip(cuts[:-1], cuts[1:])):
            end = -1 if b == E and i % 2 else b
            print(f"seed={seed} E={E} hubs={hubs} range [{a}, {end}) other={i == 0}", flush=True)
            bins, n = g.count_hub_part(hubs, a, end, other_parts=(i == 0))
            if i == 0:
                st = g.stats()
                assert st["triangles"] == -1
                wedges = st["wedges"]
            total += bins
            lt += n
    assert np.array_equal(total.astype(np.int64), want) and np.array_equal(full, want)
    assert lt == full_lt
    closed = int(want.sum() - want[OPEN].sum())
    assert int(want[OPEN].sum()) + 3 * closed == wedges


@pytest.mark.parametrize("seed", range(4))
def test_split_paths_match_oracle(gpu_ready, monkeypatch, seed):
    """The paths of graphs too big for one pass, forced on small ones: the centre pass in ranges
    whose pair types are merged (HM_HUB_TYPE_CAP just above one hyperedge's entries), rows sized by
    the upper light degrees (HM_HUB_ROWS_BY_DEGREE=0) and light-triangle launches of a few dozen
    row entries (HM_HUB_STATE_CAP=40)."""
    rng = np.random.default_rng(9300 + seed)
    kmax = [6, 40][seed % 2]
    off, nodes = random_hypergraph(rng, num_edges=int(rng.integers(150, 500)),
                                   num_nodes=int(rng.integers(kmax + 10, 120)), kmax=kmax, dup_frac=0.2)
    E = off.size - 1
    want = O.static_mochy(off, nodes)
    monkeypatch.setenv("HM_HUB_TYPE_CAP", str(E + 1))
    monkeypatch.setenv("HM_HUB_ROWS_BY_DEGREE", "0")
    monkeypatch.setenv("HM_HUB_STATE_CAP", "40")
    for tile in (0, 64):
        for h in (1, 4, 12):
            print(f"seed={seed} E={E} tile={tile} hubs={h}", flush=True)
            with hm.Graph(off, nodes) as g:
                g.set_tile(tile)
                g.set_hubs(h)
                assert np.array_equal(g.count_static().astype(np.int64), want), (tile, h)
                bins, _ = g.count_hub_part(h, 0, E // 2, other_parts=True)
                rest, _ = g.count_hub_part(h, E // 2, -1, other_parts=False)
                assert np.array_equal((bins + rest).astype(np.int64), want), (tile, h)


@pytest.mark.parametrize("seed", range(4))
def test_global_memory_histograms_match_oracle(gpu_ready, monkeypatch, seed):
    """Centres whose neighbour keys outgrow the LDS histogram are counted again with histograms in
    global memory, level by level: with the LDS table cut to 3 keys (HM_HUB_LDS_KEYS) nearly every
    centre takes that path, and with the first two global levels cut to 3 keys as well
    (HM_HUB_BIG_KEYS, odd seeds) nearly every centre reaches the last one; in one table and in
    tiles, the bins still equal the oracle's."""
    rng = np.random.default_rng(9500 + seed)
    kmax = [6, 40][seed % 2]
    off, nodes = random_hypergraph(rng, num_edges=int(rng.integers(150, 500)),
                                   num_nodes=int(rng.integers(kmax + 10, 120)), kmax=kmax, dup_frac=0.2)
    want = O.static_mochy(off, nodes)
    monkeypatch.setenv("HM_HUB_LDS_KEYS", "3")
    if seed % 2:
        monkeypatch.setenv("HM_HUB_BIG_KEYS", "3")
    for tile in (0, 64):
        for h in (1, 4, 12):
            print(f"seed={seed} E={off.size - 1} tile={tile} hubs={h}", flush=True)
            with hm.Graph(off, nodes) as g:
                g.set_tile(tile)
                g.set_hubs(h)
                assert np.array_equal(g.count_static().astype(np.int64), want), (tile, h)
                assert g.hub_stats()["big_centres"] > 0


def test_time_bounded_parts_resume_where_they_stopped(gpu_ready):
    """A part whose time has run out stops after a launch and reports the rank it reached; parts
    resumed from there, until every rank is done, sum to the count (tiled path, 1,024-centre
    launches)."""
    off, nodes = hm.generate_static(W.scaled(W.STATIC_C2, 4_000, 2_000, seed=23))
    E = off.size - 1
    with hm.Graph(off, nodes) as g:
        g.set_tile(512)
        g.set_hubs(6)
        want = g.count_static().astype(np.int64)
        total, _, done = g.count_hub_part(6, 0, -1, other_parts=True, seconds=1e-6, with_done=True)
        parts = 1
        assert 0 < done < E
        while done < E:
            bins, _, nxt = g.count_hub_part(6, done, -1, other_parts=False, seconds=1e-6, with_done=True)
            assert nxt > done
            total, done, parts = total + bins, nxt, parts + 1
        print(f"E={E}: {parts} parts", flush=True)
        assert parts >= 3 and np.array_equal(total.astype(np.int64), want)


def test_part_count_rejects_bad_arguments(gpu_ready):
    off, nodes = csr([[0, 1], [1, 2], [2, 0], [0, 3]])
    with hm.Graph(off, nodes) as g:
        for args in ((0, 0, -1), (24, 0, -1), (3, -1, -1), (3, 3, 2)):
            with pytest.raises(hm.HMotifError):
                g.count_hub_part(*args)
        with pytest.raises(hm.HMotifError):
            g.count_hub_part(3, 0, -1, seconds=-1.0)


def test_beyond_one_table_equals_triangle_path(gpu_ready):
    off, nodes = hm.generate_static(W.scaled(W.STATIC_C2, 20_000, 10_000, seed=22))
    with hm.Graph(off, nodes) as g:
        plain = g.count_static().astype(np.int64)
        for h in (6, 12, -1):
            g.set_hubs(h)
            print(f"c2 family 20k hubs={h}", flush=True)
            assert np.array_equal(g.count_static().astype(np.int64), plain), h
            assert g.hub_stats()["hubs"] > 0
    with hm.Graph(*csr([[0, 1], [1, 2], [2, 0]])) as g:
        for bad in (-2, 24):
            with pytest.raises(hm.HMotifError):
                g.set_hubs(bad)


def test_dynamic_base_full_size(gpu_ready):
    """The 500,000-hyperedge dynamic base (3.16e15 triples, its largest node in 198,573
    hyperedges): the automatic hub count against the vector two hub counts agreed on, and the wedge
    identity against the projected degrees."""
    g = json.load(open(os.path.join(GOLDEN, "static_c3_base_hub.json")))
    off, nodes = hm.generate_static(W.DYNAMIC_C3)
    assert g["params"]["num_edges"] == off.size - 1
    with hm.Graph(off, nodes) as gr:
        gr.set_hubs(-1)
        c = gr.count_static().astype(np.int64)
        st, hs = gr.stats(), gr.hub_stats()
    print(f"dynamic base: {int(c.sum())} triples, hubs {hs}", flush=True)
    assert hs["hubs"] > 0
    assert c.tolist() == g["counts30"]
    assert st["wedges"] == g["wedges"] and wedge_identity(c, st)


# ---- hub-aware deltas (hm_graph_update on a graph with hubs set) ---------------------------------
def _rows(off, nodes):
    return [list(nodes[off[e]:off[e + 1]]) for e in range(off.size - 1)]


def _batch(rng, rows, live, n_del, n_ins, num_nodes, kmax):
    dels = sorted(rng.choice(sorted(live), size=min(n_del, len(live)), replace=False).tolist())
    new_rows = []
    for _ in range(n_ins):
        if rng.random() < 0.2 and live:  # duplicate of a live hyperedge
            new_rows.append(list(rows[int(rng.choice(sorted(live)))]))
        else:
            k = int(rng.integers(1, kmax + 1))
            new_rows.append(list(rng.choice(num_nodes, size=min(k, num_nodes), replace=False)))
    return dels, new_rows


@pytest.mark.parametrize("seed", range(8))
def test_hub_deltas_match_oracle_recount(gpu_ready, seed):
    rng = np.random.default_rng(8000 + seed)
    nv, kmax = int(rng.integers(10, 80)), [4, 8, 40][seed % 3]
    off, nodes = random_hypergraph(rng, num_edges=int(rng.integers(20, 250)), num_nodes=nv, kmax=kmax)
    rows = _rows(off, nodes)
    live = set(range(len(rows)))
    hubs = [1, 3, 6, -1][seed % 4]
    with hm.Graph(off, nodes) as g:
        g.set_hubs(hubs)
        if seed % 2:
            g.set_tile(64)  # tiled light triangles
        before = O.static_mochy(*csr(rows))
        for b in range(4):
            dels, new_rows = _batch(rng, rows, live, int(rng.integers(0, 15)), int(rng.integers(0, 15)), nv, kmax)
            ids, delta = g.update(dels, *csr(new_rows))
            for x in dels:
                rows[x] = []
                live.discard(x)
            rows.extend(new_rows)
            live.update(ids.tolist())
            after = O.static_mochy(*csr(rows))
            print(f"seed={seed} hubs={hubs} batch={b} E={len(rows)}", flush=True)
            assert np.array_equal(delta, after - before)
            assert np.array_equal(g.count_static().astype(np.int64), after)
            before = after


def test_hub_deltas_equal_triangle_path_deltas(gpu_ready):
    """The dynamic workload's generator at 20k hyperedges: the same batches as hub-aware deltas
    and as k_delta deltas, equal bin for bin, and the hub-aware counts after them equal."""
    p = W.scaled(W.DYNAMIC_C3, 20_000, 8_000)
    off, nodes = hm.generate_static(p)
    with hm.Graph(off, nodes) as ga, hm.Graph(off, nodes) as gb:
        ga.set_hubs(-1)
        live = np.arange(off.size - 1)
        for b in range(1, 4):
            dels, ioff, inodes = W.dynamic_batch(live, b, params=p, inserts=280, deletes=120)
            ids_a, da = ga.update(dels, ioff, inodes)
            ids_b, db = gb.update(dels, ioff, inodes)
            print(f"batch {b}: {int(np.abs(da).sum())} changed triples, hubs {ga.hub_stats()}", flush=True)
            assert np.array_equal(ids_a, ids_b) and np.array_equal(da, db)
            live = np.setdiff1d(live, dels)
            live = np.concatenate([live, ids_a])
        assert np.array_equal(ga.count_static(), gb.count_static())


def test_dynamic_workload_full_size_first_batches(gpu_ready):
    """The full dynamic workload's first two batches on the 500,000-hyperedge base: hub-aware
    deltas equal to those of the full 100-batch run, whose recounts after batches 33, 66 and 100
    matched the running sums (tests/golden/dynamic_c3_full_hub.json)."""
    g = json.load(open(os.path.join(GOLDEN, "dynamic_c3_full_hub.json")))
    off, nodes = hm.generate_static(W.DYNAMIC_C3)
    with hm.Graph(off, nodes) as gr:
        gr.set_hubs(-1)
        live = np.arange(off.size - 1)
        for b in (1, 2):
            dels, ioff, inodes = W.dynamic_batch(live, b)
            ids, delta = gr.update(dels, ioff, inodes)
            print(f"batch {b}: sum|delta| {int(np.abs(delta).sum())}, {gr.update_stats()}", flush=True)
            assert delta.tolist() == g["deltas30"][b - 1]
            live = np.concatenate([np.setdiff1d(live, dels), ids])


def test_census_graph_counts_match_the_fixture(gpu_ready):
    """The smallest of the serving configuration's graphs above the served band (191,373
    hyperedges): the automatic hub-aware count and a two-part count (hm_graph_count_hub_part)
    equal the census vector (tests/golden/serving_c5_large_counts.json)."""
    fx = json.load(open(os.path.join(GOLDEN, "serving_c5_large_counts.json")))["counts"][0]
    off, nodes = W._serving_graph(fx["hyperedges"], fx["graph"], 5)
    assert off.size - 1 == fx["hyperedges"]
    with hm.Graph(off, nodes) as g:
        g.set_hubs(-1)
        c = g.count_static().astype(np.int64)
        hs = g.hub_stats()
        print(f"graph {fx['graph']}: {int(c.sum())} triples, {hs}", flush=True)
        assert c.tolist() == fx["counts30"]
        h = hs["hubs"]
        E = off.size - 1
        a, _ = g.count_hub_part(h, 0, E // 3, other_parts=True)
        b, _ = g.count_hub_part(h, E // 3, -1, other_parts=False)
        assert (a + b).astype(np.int64).tolist() == fx["counts30"]
