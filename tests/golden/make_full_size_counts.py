"""Recompute the exact triangle counts of the full-size graphs on the CPU.

Writes ``tests/golden/full_size_counts.json``: for each graph, the oriented
edge count and the triangle count of the C oracle (the reference's Numba CPU
exact path restated in C, cpu_exp/triangle_hll_experiment.py:543-603) on the
CSR of the oracle's numpy loader (experiment/exp_gpu.py:291-365) over the
pairs of the seeded host generators.  The GPU tests that count these graphs
at full size compare against the file.

    python tests/golden/make_full_size_counts.py [rmat23] [chunglu] [rmat20]

Host memory: about 25 GB for R-MAT scale 23 (134M raw pairs); each graph takes
about a minute on 8 cores.
"""

from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "full_size_counts.json")
GRAPHS = {
    "rmat23": "R-MAT scale 23, edge factor 16, (0.57, 0.19, 0.19, 0.05), seed 4",
    "rmat20": "R-MAT scale 20, edge factor 16, (0.57, 0.19, 0.19, 0.05), seed 5",
    "chunglu": "Chung-Lu n = 3,072,441, beta 2.1, w_max 33,000, 117M target edges, seed 2",
}


def count(name: str) -> dict:
    from oracle import cbind
    from oracle import tc_oracle as O
    from triangle_counting_hll_amd.graphgen import config_pairs

    t0 = time.perf_counter()
    pairs = config_pairs(name)
    rp, col, *_ = O.csr_from_pairs(pairs)
    del pairs
    c = cbind.exact_count(rp, col)
    return {"graph": GRAPHS[name], "n_nodes": int(len(rp) - 1), "oriented_edges": int(len(col)),
            "triangles": int(c), "max_forward_degree": int(np.diff(rp).max()),
            "seconds": round(time.perf_counter() - t0, 1), "threads": cbind.max_threads()}


def main(names):
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    for name in names:
        data[name] = count(name)
        print(name, data[name], flush=True)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
            f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1:] or ["rmat23"])
