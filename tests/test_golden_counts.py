"""The committed full-size counts reproduce from their recipe (the R-MAT scale
20 graph, about 20 s on 8 cores; the scale-23 and Chung-Lu entries take
minutes and are rebuilt by running tests/golden/make_full_size_counts.py)."""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_full_size_counts as M  # noqa: E402


def test_rmat20_entry_reproduces():
    with open(M.OUT) as f:
        data = json.load(f)
    assert set(data) == set(M.GRAPHS)
    got = M.count("rmat20")
    for k in ("n_nodes", "oriented_edges", "triangles", "max_forward_degree"):
        assert got[k] == data["rmat20"][k], k
