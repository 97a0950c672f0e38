"""The reference's own interfaces (structure.hpp, motif.hpp, motif_update.hpp, temporal_*.hpp,
re-implemented in include/compat over the C ABI) driven by a C++ program written against them
(ipdps26_amd/bin/hm_ref_check): its counts must equal the C ABI's, its delta must equal the
difference of two recounts, and the strict-inc window must give the reference driver's known
answer (temporalDynamicHMotif/src/main.cu:267-305)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, csr, random_hypergraph

import ipdps26_amd as hm
from ipdps26_amd import workloads as W

pytestmark = pytest.mark.gpu
TOOL = os.path.join(ROOT, "ipdps26_amd", "bin", "hm_ref_check")


def run_tool(tmp_path, off, nodes, delta_lines):
    csv = tmp_path / "g.csv"
    W.write_mochy_csv(str(csv), off, nodes)
    dfile = tmp_path / "d.txt"
    dfile.write_text("\n".join(delta_lines) + "\n")
    out = subprocess.run([TOOL, str(csv), str(dfile)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    res = {}
    for line in out.stdout.splitlines():
        label, _, vals = line.partition(":")
        res[label.replace(" (30 bins)", "")] = [int(x) for x in vals.split()]
    return res


@pytest.mark.parametrize("seed", range(3))
def test_reference_driver_flow(gpu_ready, tmp_path, seed):
    rng = np.random.default_rng(9500 + seed)
    off, nodes = random_hypergraph(rng, num_edges=int(rng.integers(40, 400)), num_nodes=60, kmax=8)
    rows = [list(nodes[off[e]:off[e + 1]]) for e in range(off.size - 1)]
    res = run_tool(tmp_path, off, nodes, ["A 1 3", "R 3 1"])
    with hm.Graph(off, nodes) as g:
        before = g.count_static().astype(np.int64)
    assert res["Motif counts"] == before.tolist()
    # the tool's batch: delete ids 2, 5, 7 (1-based), id 5 re-inserted with row 1, rows 3 and
    # 1 + first node of row 2 appended
    upd = [list(r) for r in rows]
    for i in (2, 5, 7):
        upd[i - 1] = []
    upd[4] = list(rows[0])
    upd.append(list(rows[2]))
    upd.append(list(rows[0]) + [rows[1][0] if rows[1] else 0])
    with hm.Graph(*csr(upd)) as g:
        after = g.count_static().astype(np.int64)
    assert res["Recount after"] == after.tolist()
    assert res["Delta counts"] == (after - before).tolist()
    # synthetic window: bin 26 = 1; delta bin 10 +1, bin 26 -1
    tc = np.array(res["Temporal strict-inc counts"])
    assert tc[26] == 1 and tc.sum() == 1
    td = np.array(res["Temporal strict-inc delta"])
    assert td[10] == 1 and td[26] == -1 and np.abs(td).sum() == 2
    # file deltas on the new window's newest layer: row 1 gains vertex 3, row 3 loses vertex 1
    assert res["Newest row 3 after file deltas"] == [2, 3]
    L = csr([[1, 2], [2, 3], [1, 3]])
    N = csr([[1, 2, 3], [2, 3], [2, 3]])
    assert res["Temporal strict-inc after file deltas"] == hm.count_strict_inc(L, L, N).astype(np.int64).tolist()
