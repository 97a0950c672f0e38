"""The committed parts of the 2M-hyperedge static count (tests/golden/static_c2_parts, made on an
MI355X by tests/golden/make_static_c2.py) combine to tests/golden/static_c2_hub.json: one part holds
the class, pair-type and light-wedge parts, the light-rank ranges tile [0, E), and the bins satisfy
sum(open) + 3 sum(closed) = sum_a C(d_a, 2) against the projected degrees."""
import glob
import json
import os

from conftest import GOLDEN

OPEN = range(20, 26)


def test_static_c2_parts_combine_to_the_fixture():
    fx = json.load(open(os.path.join(GOLDEN, "static_c2_hub.json")))
    parts = sorted((json.load(open(p)) for p in
                    glob.glob(os.path.join(GOLDEN, "static_c2_parts", f"c2_h{fx['hubs']}_part_*.json"))),
                   key=lambda r: r["begin"])
    E = fx["params"]["num_edges"]
    assert sum(r["other_parts"] for r in parts) == 1
    assert parts[0]["begin"] == 0 and parts[-1]["end"] == E
    assert all(a["end"] == b["begin"] for a, b in zip(parts, parts[1:]))
    assert all(r["hubs"] == fx["hubs"] and r["hyperedges"] == E for r in parts)
    bins = [sum(r["bins"][k] for r in parts) % (1 << 64) for k in range(30)]
    assert bins == fx["counts30"] and sum(bins) == fx["triples"]
    n_open = sum(bins[k] for k in OPEN)
    assert n_open + 3 * (sum(bins) - n_open) == fx["wedges"]
    assert sum(r["light_triangles"] for r in parts) == fx["light_triangles"]


def test_serving_c5_large_counts_hold_the_wedge_identity():
    """Every graph of the serving configuration above the served band was counted at its own size
    (tests/golden/make_c5_large_counts.py), and its bins satisfy the wedge identity."""
    import sys

    sys.path.insert(0, GOLDEN)
    import make_c5_large_counts as M

    fx = json.load(open(os.path.join(GOLDEN, "serving_c5_large_counts.json")))
    assert [(r["graph"], r["hyperedges"]) for r in fx["counts"]] == M.large_graphs()
    for r in fx["counts"]:
        c = r["counts30"]
        n_open = sum(c[k] for k in OPEN)
        assert sum(c) == r["triples"] and n_open + 3 * (sum(c) - n_open) == r["wedges"], r["graph"]


def _parts_module():
    import importlib.util
    import sys

    sys.path.insert(0, GOLDEN)
    spec = importlib.util.spec_from_file_location("make_static_c2", os.path.join(GOLDEN, "make_static_c2.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_static_c2_combine_refuses_bad_part_sets():
    """The combine step of tests/golden/make_static_c2.py refuses parts that leave a gap between
    ranges, count the other parts twice, or do not end at E, and accepts the committed set."""
    import copy

    import pytest

    M = _parts_module()
    parts = [json.load(open(p)) for p in glob.glob(os.path.join(GOLDEN, "static_c2_parts", "c2_h23_part_*.json"))]
    M.combine_hubs(copy.deepcopy(parts))  # the committed set holds
    gap = copy.deepcopy(parts)
    min(gap, key=lambda r: r["begin"])["end"] -= 1
    twice = copy.deepcopy(parts)
    for r in twice:
        r["other_parts"] = True
    short = copy.deepcopy(parts)
    max(short, key=lambda r: r["begin"])["end"] -= 1
    wrong = copy.deepcopy(parts)
    wrong[0]["bins"][3] += 1
    for bad in (gap, twice, short, wrong):
        with pytest.raises(AssertionError):
            M.combine_hubs(bad)
