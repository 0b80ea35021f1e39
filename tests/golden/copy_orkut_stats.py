"""Copies the reference's published Orkut degree-statistics outputs into this
directory as fixtures (data only: the degree histogram CSV, the log2-bucket CSV
and the summary JSON written by experiment/gpu_graph_stats.py:186-320 on
com-orkut.ungraph.txt).  Run once in a checkout that has the reference at
/root/reference; the tests read only the copies."""

import os
import shutil

SRC = "/root/reference/results/orkut_graph_stats"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "orkut_graph_stats")

if __name__ == "__main__":
    os.makedirs(DST, exist_ok=True)
    for name in ("degree_distribution.csv", "degree_log2_buckets.csv", "graph_stats.json"):
        shutil.copyfile(os.path.join(SRC, name), os.path.join(DST, name))
        print("copied", name)
