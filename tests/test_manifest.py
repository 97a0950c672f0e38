"""The named workloads are regenerated bit-identically (SHA-256 manifest written by
tests/golden/make_manifest.py): the seeded generator is the inputs' checkpoint."""
import json
import os

from conftest import GOLDEN

import sys
sys.path.insert(0, GOLDEN)
import make_manifest  # noqa: E402


def test_workloads_match_manifest():
    with open(os.path.join(GOLDEN, "workloads_manifest.json")) as f:
        want = json.load(f)
    got = make_manifest.entries()
    assert sorted(got) == sorted(want)
    for k in want:
        assert got[k] == want[k], k
