"""Serving: every request's result is the same alone (plain, one request at a time) and in the mix
(served, concurrent slots, coalesced counts), admission respects the HBM budget."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import csr

import ipdps26_amd as hm
from ipdps26_amd import serving, workloads as W

pytestmark = pytest.mark.gpu


def small_stream(n_graphs=10, n_req=120, rate=0.0):
    graphs = W.serving_graphs(n_graphs, lo=200, hi=1500, seed=7)
    reqs = W.serving_requests(graphs, n_req, seed=77, rate_per_s=rate)
    return graphs, reqs


def test_served_equals_plain(gpu_ready):
    graphs, reqs = small_stream()
    plain = serving.run_plain(graphs, reqs, realtime=False)
    served = serving.run_served(graphs, reqs, realtime=False, slots=4)
    assert len(plain["results"]) == len(served["results"]) == len(reqs)
    for k, (a, b) in enumerate(zip(plain["results"], served["results"])):
        assert np.array_equal(a, b), f"request {k} ({reqs[k]['kind']})"
    # a count that precedes every update of its graph is the oracle's count of the input graph
    touched, checked = set(), 0
    for q, res in zip(reqs, served["results"]):
        if q["kind"] == "update":
            touched.add(q["graph"])
        elif q["graph"] not in touched:
            assert np.array_equal(res, O.static_mochy(*graphs[q["graph"]]))
            checked += 1
    assert checked > 0


def test_served_with_arrivals_and_one_slot(gpu_ready):
    graphs, reqs = small_stream(n_graphs=6, n_req=60, rate=2000.0)
    plain = serving.run_plain(graphs, reqs, realtime=True)
    served = serving.run_served(graphs, reqs, realtime=True, slots=1)
    for a, b in zip(plain["results"], served["results"]):
        assert np.array_equal(a, b)


def test_admission_budget(gpu_ready):
    off, nodes = hm.generate_static(W.scaled(W.STATIC_C2, 2000, 1000))
    with hm.Server(hbm_budget_bytes=256 << 10, slots=2) as srv:
        with pytest.raises(hm.HMotifError) as ei:
            srv.add_graph(off, nodes)
        assert ei.value.status == 2
    with hm.Server(hbm_budget_bytes=1 << 30, slots=2) as srv:
        gid = srv.add_graph(off, nodes)
        st = srv.stats()
        assert st["graphs"] == 1 and 0 < st["reserved_bytes"] <= st["budget_bytes"]
        c = srv.wait(srv.submit_count(gid))["counts"]
        with hm.Graph(off, nodes) as g:
            assert np.array_equal(c, g.count_static().astype(np.int64))
        srv.remove_graph(gid)
        assert srv.stats()["reserved_bytes"] == 0
        with pytest.raises(hm.HMotifError):
            srv.submit_count(gid)


def test_budget_charges_growth(gpu_ready):
    """Buffers a graph grows after its admission (the store, the update's projection rows, count
    scratch) are charged: reserved_bytes rises to what the graph holds; once that is over the
    budget, an update fails with HM_ERR_OVER_BUDGET and leaves the graph unchanged, counts still
    run, and removing the graph releases everything."""
    off, nodes = hm.generate_static(W.scaled(W.STATIC_C2, 2000, 1000))
    with hm.Server(1 << 30, 1) as probe:
        probe.add_graph(off, nodes)
        est = probe.stats()["reserved_bytes"]
    budget = est + (1 << 20)
    ins_off, ins_nodes = hm.generate_static(W.scaled(W.STATIC_C2, 6000, 1000, seed=99))
    with hm.Server(budget, 1) as srv, hm.Graph(off, nodes) as g:
        gid = srv.add_graph(off, nodes)
        r = srv.wait(srv.submit_update(gid, [], ins_off, ins_nodes))
        _, d = g.update([], ins_off, ins_nodes)
        assert np.array_equal(r["counts"], d)
        held = srv.stats()["reserved_bytes"]
        print(f"estimate {est} B, budget {budget} B, held after the update {held} B", flush=True)
        assert held > budget
        with pytest.raises(hm.HMotifError) as ei:
            srv.wait(srv.submit_update(gid, [0], np.zeros(1, np.int64), np.zeros(0, np.int32)))
        assert ei.value.status == 2
        c = srv.wait(srv.submit_count(gid))["counts"]
        assert np.array_equal(c, g.count_static().astype(np.int64))  # the failed update changed nothing
        srv.remove_graph(gid)
        assert srv.stats()["reserved_bytes"] == 0


def test_poll_until_done(gpu_ready):
    graphs = W.serving_graphs(3, lo=300, hi=600, seed=8)
    with hm.Server(1 << 30, 2) as srv:
        gids = [srv.add_graph(off, nodes) for off, nodes in graphs]
        tickets = [srv.submit_count(g) for g in gids]
        got = {}
        import time
        t_end = time.time() + 60
        while len(got) < len(tickets) and time.time() < t_end:
            for t in tickets:
                if t not in got:
                    r = srv.poll(t)
                    if r is not None:
                        got[t] = r
            time.sleep(0.001)
        assert len(got) == len(tickets)
        for (off, nodes), t in zip(graphs, tickets):
            with hm.Graph(off, nodes) as g:
                assert np.array_equal(got[t]["counts"], g.count_static().astype(np.int64))
        with pytest.raises(hm.HMotifError):
            srv.poll(tickets[0])  # consumed


def test_add_and_remove_graphs_while_requests_run(gpu_ready):
    graphs = W.serving_graphs(12, lo=300, hi=1200, seed=9)
    with hm.Server(1 << 30, 4) as srv:
        first = srv.add_graph(*graphs[0])
        tickets = [srv.submit_count(first) for _ in range(6)]
        more = [srv.add_graph(off, nodes) for off, nodes in graphs[1:]]  # while those run
        tickets += [srv.submit_count(g) for g in more]
        results = [srv.wait(t)["counts"] for t in tickets]
        for g in more[:5]:
            srv.remove_graph(g)
        extra = srv.wait(srv.submit_count(more[5]))["counts"]
    expect = []
    for off, nodes in graphs:
        with hm.Graph(off, nodes) as g:
            expect.append(g.count_static().astype(np.int64))
    for r in results[:6]:
        assert np.array_equal(r, expect[0])
    for k, r in enumerate(results[6:]):
        assert np.array_equal(r, expect[k + 1])
    assert np.array_equal(extra, expect[6])


# ---- kept counts (maintain_counts = 1) ----------------------------------------------------------
def test_maintained_counts_equal_plain(gpu_ready):
    """Counts answered from kept counts (previous count + exact update deltas) equal the plain
    path's recount request by request."""
    graphs, reqs = small_stream(n_graphs=8, n_req=160)
    plain = serving.run_plain(graphs, reqs, realtime=False)
    kept = serving.run_served(graphs, reqs, realtime=False, slots=4, maintain=True)
    for k, (a, b) in enumerate(zip(plain["results"], kept["results"])):
        assert np.array_equal(a, b), f"request {k} ({reqs[k]['kind']})"
    n_counts = sum(q["kind"] == "count" for q in reqs)
    assert 0 < kept["maintained"] <= n_counts


def test_maintained_counts_start_and_survive_a_failed_update(gpu_ready):
    """No kept counts before the first count (an update alone keeps none); a failed update drops
    them, so the next count runs on the device again; every answer equals a fresh count."""
    (off, nodes), = W.serving_graphs(1, lo=800, hi=800, seed=11)
    rng = np.random.default_rng(11)
    with hm.Server(1 << 30, 2, maintain_counts=True) as srv, hm.Graph(off, nodes) as g:
        gid = srv.add_graph(off, nodes)

        def both_update(dels, ins):
            o, n = csr(ins)
            r = srv.wait(srv.submit_update(gid, dels, o, n))
            _, d = g.update(dels, o, n)
            assert np.array_equal(r["counts"], d)

        both_update([0, 5], [[1, 2, 3], [2, 3, 4, 5]])
        assert srv.stats()["maintained"] == 0
        c = srv.wait(srv.submit_count(gid))["counts"]
        assert np.array_equal(c, g.count_static().astype(np.int64))
        assert srv.stats()["maintained"] == 0  # the first count ran on the device
        live = [x for x in range(off.size - 1) if x not in (0, 5)]
        for k in range(3):
            both_update(sorted(rng.choice(live, size=3, replace=False).tolist()) if k == 0 else [],
                        [sorted(rng.choice(400, size=4, replace=False).tolist())])
            c = srv.wait(srv.submit_count(gid))["counts"]
            assert np.array_equal(c, g.count_static().astype(np.int64))
        assert srv.stats()["maintained"] == 3
        with pytest.raises(hm.HMotifError):  # id 0 was deleted: the update fails
            srv.wait(srv.submit_update(gid, [0], np.zeros(1, np.int64), np.zeros(0, np.int32)))
        c = srv.wait(srv.submit_count(gid))["counts"]
        assert np.array_equal(c, g.count_static().astype(np.int64))
        assert srv.stats()["maintained"] == 3  # recounted after the failure
        c = srv.wait(srv.submit_count(gid))["counts"]
        assert srv.stats()["maintained"] == 4
        assert np.array_equal(c, g.count_static().astype(np.int64))


# ---- batched counts (batch_graphs > 1) ----------------------------------------------------------
@pytest.mark.parametrize("maintain", [False, True])
def test_batched_counts_equal_plain(gpu_ready, maintain):
    """Counts of several ready graphs run in one batched launch sequence; every result equals the
    plain path's, with and without kept counts."""
    graphs, reqs = small_stream(n_graphs=12, n_req=200)
    plain = serving.run_plain(graphs, reqs, realtime=False)
    served = serving.run_served(graphs, reqs, realtime=False, slots=2, maintain=maintain,
                                batch_graphs=8)
    for k, (a, b) in enumerate(zip(plain["results"], served["results"])):
        assert np.array_equal(a, b), f"request {k} ({reqs[k]['kind']})"
    if not maintain:  # all submitted at once: several small graphs are ready with a count at a time
        assert served["batched_counts"] > 0 and served["batch_launches"] > 0


def test_batched_counts_with_graph_removal(gpu_ready):
    """Graphs removed while batched counts of other graphs are queued: the remaining graphs'
    counts stay exact, removal waits for a graph's own requests."""
    graphs = W.serving_graphs(10, lo=150, hi=600, seed=31)
    expect = []
    for off, nodes in graphs:
        with hm.Graph(off, nodes) as g:
            expect.append(g.count_static().astype(np.int64))
    with hm.Server(1 << 30, 2, batch_graphs=8) as srv:
        gids = [srv.add_graph(off, nodes) for off, nodes in graphs]
        for t in [srv.submit_count(g) for g in gids]:  # first counts: alone, they time the graphs
            srv.wait(t)
        tickets = [(k, srv.submit_count(g)) for _ in range(4) for k, g in enumerate(gids)]
        for k in (1, 4, 7):
            srv.remove_graph(gids[k])  # waits for graph k's queued counts
        for k, t in tickets:
            assert np.array_equal(srv.wait(t)["counts"], expect[k]), k
        again = [(k, srv.submit_count(gids[k])) for k in range(10) if k not in (1, 4, 7)]
        for k, t in again:
            assert np.array_equal(srv.wait(t)["counts"], expect[k]), k
        st = srv.stats()
        assert st["graphs"] == 7 and st["batch_launches"] > 0


# ---- captured updates (capture_updates = 1) -----------------------------------------------------
@pytest.mark.parametrize("maintain", [False, True])
def test_captured_updates_equal_plain(gpu_ready, maintain):
    """Every graph's updates run as one captured hipGraph each while other slots count, capture
    and allocate concurrently: every result equals the plain path's, request by request."""
    graphs, reqs = small_stream(n_graphs=10, n_req=200)
    plain = serving.run_plain(graphs, reqs, realtime=False)
    served = serving.run_served(graphs, reqs, realtime=False, slots=4, maintain=maintain, capture=True)
    for k, (a, b) in enumerate(zip(plain["results"], served["results"])):
        assert np.array_equal(a, b), f"request {k} ({reqs[k]['kind']})"


def test_server_config_rejects_bad_capture_flag(gpu_ready):
    import ctypes

    from ipdps26_amd import _ServerConfig, lib

    h = ctypes.c_void_p()
    cfg = _ServerConfig(1 << 30, 2, 0, 0, 2, 0)
    assert lib().hm_server_create(ctypes.byref(cfg), ctypes.byref(h)) != 0
    assert "capture_updates" in lib().hm_last_error().decode()
    cfg = _ServerConfig(1 << 30, 2, 0, 0, 0, -1)
    assert lib().hm_server_create(ctypes.byref(cfg), ctypes.byref(h)) != 0
    assert "hub_min_edges" in lib().hm_last_error().decode()


# ---- hub-aware graphs (hub_min_edges) -----------------------------------------------------------
@pytest.mark.parametrize("maintain", [False, True])
def test_hub_graphs_served_equal_plain_by_triangles(gpu_ready, maintain):
    """Graphs of 1,500 hyperedges or more counted and updated hub-aware by the server (and by the
    plain path in the second run) while smaller graphs go by triangles: every result equals the
    plain path's by triangles, request by request."""
    graphs = W.serving_graphs(8, lo=400, hi=4000, seed=41)
    reqs = W.serving_requests(graphs, 160, seed=410)
    assert any(off.size - 1 >= 1500 for off, _ in graphs) and any(off.size - 1 < 1500 for off, _ in graphs)
    plain = serving.run_plain(graphs, reqs, realtime=False)
    served = serving.run_served(graphs, reqs, realtime=False, slots=4, maintain=maintain, batch_graphs=4,
                                hub_min_edges=1500)
    plain_hub = serving.run_plain(graphs, reqs, realtime=False, hub_min_edges=1500)
    for k, (a, b, c) in enumerate(zip(plain["results"], served["results"], plain_hub["results"])):
        assert np.array_equal(a, b) and np.array_equal(a, c), f"request {k} ({reqs[k]['kind']})"
