"""GPU parity of the tiled static count (graphs whose hyperedge ids exceed one LDS table, or any
graph with a forced tile) against the oracle, the reference's golden vectors and the single-table
count. Integer counts: bit-exact."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import load_golden, random_hypergraph
from test_static_gpu import golden_input

import ipdps26_amd as hm

pytestmark = pytest.mark.gpu


def count(off, nodes, tile=0):
    with hm.Graph(off, nodes) as g:
        g.set_tile(tile)
        c = g.count_static().astype(np.int64)
        return c, g.stats()


@pytest.mark.parametrize("seed", range(12))
def test_forced_tiles_match_oracle(gpu_ready, seed):
    rng = np.random.default_rng(5000 + seed)
    kmax = [4, 12, 30, 45][seed % 4]  # 45: 64-bit masks
    tile = [64, 97, 256, 1000][(seed // 4) % 4]
    off, nodes = random_hypergraph(rng, num_edges=int(rng.integers(50, 1500)),
                                   num_nodes=int(rng.integers(kmax, 400)), kmax=kmax)
    print(f"seed={seed} E={off.size - 1} tile={tile}", flush=True)
    c, _ = count(off, nodes, tile)
    assert np.array_equal(c, O.static_mochy(off, nodes))


@pytest.mark.parametrize("seed", range(3))
def test_forced_tiles_brute_force(gpu_ready, seed):
    rng = np.random.default_rng(6000 + seed)
    off, nodes = random_hypergraph(rng, num_edges=60, num_nodes=15, kmax=6, dup_frac=0.3)
    for tile in (1, 7, 64):
        c, _ = count(off, nodes, tile)
        assert np.array_equal(c, O.static_brute(off, nodes)), tile


@pytest.mark.parametrize("name", ["dblp", "c1_family_e800", "c2_family_e3000", "static_c1",
                                  "c5_family_e10000"])
def test_goldens_with_tiles(gpu_ready, name):
    gs = [x for x in load_golden("static") if x["name"] == name]
    if not gs:
        pytest.skip(f"golden {name} not generated")
    off, nodes = golden_input(gs[0])
    single, st1 = count(off, nodes, 0)
    tile = max(64, (off.size - 1) // 7)
    print(f"golden={name} E={off.size - 1} tile={tile}", flush=True)
    tiled, st2 = count(off, nodes, tile)
    assert hm.printed_motifs(tiled) == gs[0]["printed_motifs"]
    assert np.array_equal(tiled, single)
    for k in ("proj_pairs", "max_proj_degree", "oriented_row_sum", "wedges", "triangles"):
        assert st1[k] == st2[k], k


def test_beyond_one_table_matches_oracle(gpu_ready):
    # 50,000 hyperedge ids: more than one LDS table holds, so the automatic path tiles
    p = hm.GenParams(num_nodes=500_000, num_edges=50_000, alpha=2.8, kmin=2, kmax=12, zipf_s=0.7,
                     seed=12)
    off, nodes = hm.generate_static(p)
    c, st = count(off, nodes, 0)
    ref = O.static_mochy(off, nodes)
    assert np.array_equal(c, ref)
    deg = O.projection_degrees(off, nodes)
    assert st["proj_pairs"] == int(deg.sum()) and st["max_proj_degree"] == int(deg.max())
    c2, _ = count(off, nodes, 4093)  # another tiling, same counts
    assert np.array_equal(c2, c)


def test_beyond_one_table_wide_masks(gpu_ready):
    # hyperedges of up to 40 nodes (64-bit masks), 40,000 ids; tile independence and the exact
    # identity  open + 3 closed = sum_a C(d_a, 2)  with d_a from the oracle's projection
    p = hm.GenParams(num_nodes=100_000, num_edges=40_000, alpha=2.5, kmin=2, kmax=40, zipf_s=0.8,
                     seed=12)
    off, nodes = hm.generate_static(p)
    c, st = count(off, nodes, 0)
    c2, _ = count(off, nodes, 5000)
    assert np.array_equal(c, c2)
    deg = O.projection_degrees(off, nodes).astype(np.int64)
    open_ = int(c[20:26].sum())
    closed = int(c.sum()) - open_
    assert open_ + 3 * closed == int((deg * (deg - 1) // 2).sum())
    assert closed == st["triangles"] and st["proj_pairs"] == int(deg.sum())


@pytest.mark.parametrize("E", [28_672, 28_673])
def test_single_table_boundary(gpu_ready, E):
    # 28,672 ids (u32 masks) is the largest single LDS table next to the 20 KB of histograms; one
    # more id switches to tiles
    p = hm.GenParams(num_nodes=300_000, num_edges=E, alpha=2.8, kmin=2, kmax=12, zipf_s=0.7, seed=13)
    off, nodes = hm.generate_static(p)
    c, _ = count(off, nodes, 0)
    assert np.array_equal(c, O.static_mochy(off, nodes))


def test_tile_argument_checks(gpu_ready):
    off, nodes = random_hypergraph(np.random.default_rng(1), 20, 30, 5)
    with hm.Graph(off, nodes) as g:
        with pytest.raises(hm.HMotifError):
            g.set_tile(-1)
        g.set_tile(1 << 24)  # larger than LDS: rejected at count time
        with pytest.raises(hm.HMotifError) as ei:
            g.count_static()
        assert ei.value.status == 3
