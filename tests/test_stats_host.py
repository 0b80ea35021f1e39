"""Degree-statistics host logic pinned by the reference's published Orkut outputs
(tests/golden/orkut_graph_stats, written by experiment/gpu_graph_stats.py)."""

import csv
import json
import os

import numpy as np

from triangle_counting_hll_amd import graph_stats

GOLD = os.path.join(os.path.dirname(__file__), "golden", "orkut_graph_stats")


def orkut_hist():
    rows = list(csv.reader(open(os.path.join(GOLD, "degree_distribution.csv"))))[1:]
    deg = np.array([int(r[0]) for r in rows])
    hist = np.zeros(deg.max() + 1, dtype=np.int64)
    hist[deg] = [int(r[1]) for r in rows]
    return hist


def test_summary_formulas_reproduce_orkut_json():
    want = json.load(open(os.path.join(GOLD, "graph_stats.json")))
    hist = orkut_hist()
    d = np.arange(len(hist), dtype=object)
    active = int(hist.sum())
    s1 = int((d * hist).sum())
    s2 = int((d * d * hist).sum())
    got = graph_stats.summary_from_sums(active, s1, s2, len(hist) - 1,
                                        int(np.nonzero(hist)[0][0]))
    for k in ("active_nodes", "degree_sum", "avg_degree", "min_degree", "max_degree",
              "std_degree", "density_active_nodes"):
        assert got[k] == want[k], k
    assert int(hist[1]) == want["degree_1_nodes"]


def test_csv_writers_reproduce_orkut_files(tmp_path):
    hist = orkut_hist()
    active = int(hist.sum())
    nz = np.nonzero(hist)[0]
    log_hist = np.zeros(graph_stats.n_log_buckets_for(len(hist) - 1), dtype=np.int64)
    np.add.at(log_hist, np.floor(np.log2(nz)).astype(int), hist[nz])
    graph_stats.write_degree_csv(str(tmp_path / "d.csv"), hist, active)
    graph_stats.write_log_csv(str(tmp_path / "l.csv"), log_hist, active)
    for mine, ref in (("d.csv", "degree_distribution.csv"), ("l.csv", "degree_log2_buckets.csv")):
        a = (tmp_path / mine).read_bytes().replace(b"\r\n", b"\n")
        b = open(os.path.join(GOLD, ref), "rb").read().replace(b"\r\n", b"\n")
        assert a == b, mine
