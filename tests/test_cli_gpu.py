"""The command-line tools keep the reference programs' process contracts (arguments, stdout and
result-file formats) and give the reference's numbers: hm_exact_par vs MoCHy main_exact_par,
hm_thymeP vs THyMe+ main_thymeP."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, load_golden

from ipdps26_amd import workloads as W

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "ipdps26_amd", "bin")


def test_hm_exact_par_on_dblp(gpu_ready):
    g = [x for x in load_golden("static") if x["name"] == "dblp"][0]
    out = subprocess.run([os.path.join(BIN, "hm_exact_par"), "8", os.path.join(GOLDEN, g["file"])],
                         capture_output=True, text=True, check=True, timeout=120).stdout
    motifs = [int(m.group(2)) for m in re.finditer(r"^motif (\d+): (\d+)$", out, re.M)]
    assert motifs == g["printed_motifs"]
    assert re.search(r"^TOTAL_RUNTIME_SECONDS [0-9.]+", out, re.M)
    assert "# of hyperedges: 5596" in out


def test_hm_exact_par_default_file_and_usage(gpu_ready, tmp_path):
    r = subprocess.run([os.path.join(BIN, "hm_exact_par")], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr
    # default graph file name, resolved in the working directory as MoCHy does
    (tmp_path / "dblp_graph.txt").write_text(open(os.path.join(GOLDEN, "dblp_graph.txt")).read())
    out = subprocess.run([os.path.join(BIN, "hm_exact_par"), "4"], cwd=tmp_path, capture_output=True,
                         text=True, check=True, timeout=120).stdout
    assert out.count("motif ") == 26


def test_hm_thymeP_result_file(gpu_ready, tmp_path):
    g = [x for x in load_golden("temporal") if x["name"] == "enron_distinct_d100000"][0]
    off, nodes, _ = W.read_thyme_tsv(os.path.join(GOLDEN, g["file"]))
    times = (np.arange(off.size - 1, dtype=np.int64) + 1) * 1000
    tsv, res = tmp_path / "enron.tsv", tmp_path / "res.txt"
    W.write_thyme_tsv(str(tsv), off, nodes, times)
    out = subprocess.run([os.path.join(BIN, "hm_thymeP"), "x", str(g["delta"]), "8", str(tsv), str(res)],
                         capture_output=True, text=True, check=True, timeout=120).stdout
    assert "TOTAL_RUNTIME_SECONDS" in out
    lines = res.read_text().split("\n")
    float(lines[0])
    assert [int(l.split("\t")[0]) for l in lines[1:97]] == list(range(1, 97))
    assert [int(l.split("\t")[1]) for l in lines[1:97]] == g["counts96"]
