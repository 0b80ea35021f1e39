"""Degree statistics on the GPU against the reference's published Orkut outputs
and against numpy on generated edge lists."""

import csv
import json
import os

import numpy as np
import pytest

from triangle_counting_hll_amd import graph_stats, graphgen

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "orkut_graph_stats")


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.reader(f))


def test_orkut_degrees_reproduce_published_stats(device, tmp_path):
    # a degree array with Orkut's histogram (tests/golden, from the reference's
    # results/orkut_graph_stats), spread over max_node_id + 1 slots
    want = json.load(open(os.path.join(GOLD, "graph_stats.json")))
    rows = _rows(os.path.join(GOLD, "degree_distribution.csv"))[1:]
    deg = np.repeat([int(r[0]) for r in rows], [int(r[1]) for r in rows]).astype(np.int32)
    n_slots = want["max_node_id"] + 1
    slots = np.zeros(n_slots, np.int32)
    slots[np.random.default_rng(0).permutation(n_slots)[:len(deg)]] = deg
    summ = graph_stats.degree_summary_gpu(slots)
    for k in ("active_nodes", "degree_sum", "avg_degree", "min_degree", "max_degree",
              "std_degree", "density_active_nodes"):
        assert summ[k] == want[k], k
    hist, log_hist = graph_stats.degree_histograms_gpu(slots, summ["max_degree"])
    graph_stats.write_degree_csv(str(tmp_path / "d.csv"), hist, summ["active_nodes"])
    graph_stats.write_log_csv(str(tmp_path / "l.csv"), log_hist, summ["active_nodes"])
    assert _rows(tmp_path / "d.csv") == _rows(os.path.join(GOLD, "degree_distribution.csv"))
    assert _rows(tmp_path / "l.csv") == _rows(os.path.join(GOLD, "degree_log2_buckets.csv"))
    top = graph_stats.top_degrees_gpu(slots, 20)
    assert [t["degree"] for t in top] == [t["degree"] for t in want["top_degrees"]]


@pytest.mark.parametrize("dedup", [False, True])
def test_analyze_matches_numpy(device, tmp_path, dedup):
    pairs = graphgen.rmat(13, 8, seed=11)
    pairs = np.concatenate([pairs, pairs[:50, ::-1], [[9, 9]]])
    path = tmp_path / "g.txt"
    path.write_bytes(graphgen.edge_list_text(pairs, header="# test\n"))
    res = graph_stats.analyze_graph_gpu(str(path), str(tmp_path / "out"), dedup=dedup, top_k=5)
    # the reference's steps in numpy (gpu_graph_stats.py:59-86, 89-140)
    e = pairs[pairs[:, 0] != pairs[:, 1]]
    e = np.sort(e, axis=1)
    if dedup:
        e = np.unique(e, axis=0)
    n_slots = int(e.max()) + 1
    deg = np.zeros(n_slots, np.int64)
    np.add.at(deg, e[:, 0], 1)
    np.add.at(deg, e[:, 1], 1)
    pos = deg[deg > 0]
    assert res["raw_edges"] == int((pairs[:, 0] != pairs[:, 1]).sum())
    assert res["used_edges"] == len(e) and res["max_node_id"] == n_slots - 1
    assert res["active_nodes"] == len(pos) and res["degree_sum"] == int(pos.sum())
    assert res["max_degree"] == int(pos.max()) and res["min_degree"] == int(pos.min())
    assert res["degree_1_nodes"] == int((deg == 1).sum())
    order = sorted(range(n_slots), key=lambda i: (-deg[i], i))[:5]
    assert res["top_degrees"] == [{"node": i, "degree": int(deg[i])} for i in order]
    rows = _rows(tmp_path / "out" / "degree_distribution.csv")[1:]
    vals, cnts = np.unique(pos, return_counts=True)
    assert [(int(r[0]), int(r[1])) for r in rows] == list(zip(vals.tolist(), cnts.tolist()))
    assert json.load(open(tmp_path / "out" / "graph_stats.json"))["degree_sum"] == res["degree_sum"]


def test_out_of_range_ids_are_rejected(device):
    with pytest.raises(ValueError, match="outside"):
        graph_stats.degree_count_gpu(np.array([[0, 1], [2, 7]], np.int32), 5)
    with pytest.raises(ValueError, match="histogram"):
        graph_stats.degree_histograms_gpu(np.array([1, 9], np.int32), 4)
