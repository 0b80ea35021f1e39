"""Copies published result files of the reference into this directory as
fixtures (data only):

* orkut_graph_stats/: the degree histogram CSV, the log2-bucket CSV and the
  summary JSON written by experiment/gpu_graph_stats.py:186-320 on
  com-orkut.ungraph.txt;
* exp_gpu_schema/ca-hepth_exp_gpu_full.json: one result file of
  experiment/exp_gpu.py:641-933 (run_exp_gpu), the schema the experiment
  harness writes;
* prototype_csv/gpu_timing_results.csv: the timing table written by
  experiment1_gpu/gpu_experiment_fixed.py:664-697 (run_gpu_experiment).

Run once in a checkout that has the reference at /root/reference; the tests
read only the copies."""

import os
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
FILES = [
    ("results/orkut_graph_stats/degree_distribution.csv",
     "orkut_graph_stats/degree_distribution.csv"),
    ("results/orkut_graph_stats/degree_log2_buckets.csv",
     "orkut_graph_stats/degree_log2_buckets.csv"),
    ("results/orkut_graph_stats/graph_stats.json", "orkut_graph_stats/graph_stats.json"),
    ("results/ca-hepth/exp_gpu_full.json", "exp_gpu_schema/ca-hepth_exp_gpu_full.json"),
    ("results_gpu_fixed/gpu_timing_results.csv", "prototype_csv/gpu_timing_results.csv"),
]

if __name__ == "__main__":
    for src, dst in FILES:
        os.makedirs(os.path.dirname(os.path.join(HERE, dst)), exist_ok=True)
        shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, dst))
        print("copied", src, "->", dst)
