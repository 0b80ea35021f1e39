"""Graph structure and degree distribution on the device
(drop-in for ``experiment/gpu_graph_stats.py``).

``analyze_graph_gpu(graph_path, output_dir, dedup=False, top_k=20)`` writes the
reference's three outputs -- ``degree_distribution.csv``,
``degree_log2_buckets.csv`` and ``graph_stats.json`` -- with the same columns,
keys and formulas (gpu_graph_stats.py:143-320).  The edge list is parsed on the
device (``ingest``), degrees, summary sums and histograms come from HIP kernels
(``tc_degree_count``, ``tc_degree_summary``, ``tc_degree_histogram``); the
floating-point summary is formed on the host from the kernels' integer sums
with the reference's expressions.
"""

from __future__ import annotations

import csv
import json
import math
import os
import time

import numpy as np

from . import _lib, gpu_graph, ingest

COMMENT_CHARS = "#%"  # experiment/gpu_graph_stats.py:32


@_lib.on_stream
def degree_count_gpu(pairs, n_slots: int, stream=None):
    """Undirected degree per id (int32[n_slots]) of raw pairs ``int32[m, 2]`` on the device."""
    t = _lib.require_device()
    p = _lib.as_device(pairs, t.int32, "pairs")
    deg = t.zeros(max(n_slots, 0), dtype=t.int32, device="cuda")
    m = p.shape[0] if p.ndim == 2 else 0
    base = _lib.ptr(p)
    _lib.check(_lib.load().tc_degree_count(base, base + 4 if m else None, 2, m, _lib.ptr(deg),
                                           n_slots, _lib.stream_ptr(stream)))
    return deg


@_lib.on_stream
def degree_summary_gpu(degree, stream=None) -> dict:
    """Summary of the positive degrees, with the reference's formulas
    (gpu_graph_stats.py:243-257)."""
    t = _lib.require_device()
    d = _lib.as_device(degree, t.int32, "degree")
    sums = t.empty(3, dtype=t.int64, device="cuda")
    minmax = t.empty(2, dtype=t.int32, device="cuda")
    _lib.check(_lib.load().tc_degree_summary(_lib.ptr(d), d.numel(), _lib.ptr(sums),
                                             _lib.ptr(minmax), _lib.stream_ptr(stream)))
    active, degree_sum, degree_sumsq = (int(x) for x in sums.cpu().tolist())
    max_degree, min_pos = (int(x) for x in minmax.cpu().tolist())
    return summary_from_sums(active, degree_sum, degree_sumsq, max_degree, min_pos)


def summary_from_sums(active: int, degree_sum: int, degree_sumsq: int, max_degree: int,
                      min_positive: int) -> dict:
    """The reference's summary expressions (gpu_graph_stats.py:248-257) on integer sums."""
    min_degree = min_positive if active else 0
    mean_degree = degree_sum / active if active else 0.0
    variance = (degree_sumsq / active - mean_degree * mean_degree) if active else 0.0
    std_degree = math.sqrt(max(0.0, variance))
    density = degree_sum / (active * (active - 1)) if active > 1 else 0.0
    return {"active_nodes": active, "degree_sum": degree_sum, "degree_sumsq": degree_sumsq,
            "max_degree": max_degree, "min_degree": min_degree, "avg_degree": mean_degree,
            "std_degree": std_degree, "density_active_nodes": density}


def n_log_buckets_for(max_degree: int) -> int:
    """``max(1, floor(log2(max_degree)) + 1)`` (gpu_graph_stats.py:267)."""
    return max(1, int(math.floor(math.log2(max_degree))) + 1) if max_degree > 0 else 1


@_lib.on_stream
def degree_histograms_gpu(degree, max_degree: int, stream=None):
    """(hist int64[max_degree + 1], log_hist int64[n_log_buckets]) as host arrays."""
    t = _lib.require_device()
    d = _lib.as_device(degree, t.int32, "degree")
    n_log = n_log_buckets_for(max_degree)
    hist = t.empty(max_degree + 1, dtype=t.int64, device="cuda")
    log_hist = t.empty(n_log, dtype=t.int64, device="cuda")
    _lib.check(_lib.load().tc_degree_histogram(_lib.ptr(d), d.numel(), _lib.ptr(hist),
                                               hist.numel(), _lib.ptr(log_hist), n_log,
                                               _lib.stream_ptr(stream)))
    return hist.cpu().numpy(), log_hist.cpu().numpy()


def write_degree_csv(path: str, hist: np.ndarray, active_nodes: int) -> None:
    """degree,node_count,fraction rows for every non-empty degree
    (gpu_graph_stats.py:151-158)."""
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["degree", "node_count", "fraction"])
        for d in np.nonzero(hist)[0]:
            cnt = int(hist[d])
            w.writerow([int(d), cnt, cnt / active_nodes if active_nodes else 0.0])


def write_log_csv(path: str, log_hist: np.ndarray, active_nodes: int) -> None:
    """bucket,degree_low,degree_high,node_count,fraction rows for every non-empty
    log2 bucket (gpu_graph_stats.py:161-180)."""
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["bucket", "degree_low", "degree_high", "node_count", "fraction"])
        for b, cnt in enumerate(log_hist):
            cnt = int(cnt)
            if cnt == 0:
                continue
            low, high = 1 << b, (1 << (b + 1)) - 1
            w.writerow([f"[{low},{high}]", low, high, cnt,
                        cnt / active_nodes if active_nodes else 0.0])


def top_degrees_gpu(degree, top_k: int) -> list:
    """The ``top_k`` highest-degree ids, highest first; equal degrees by ascending id
    (the reference's argpartition/argsort leaves that order unspecified)."""
    t = _lib.require_device()
    d = _lib.as_device(degree, t.int32, "degree").long()
    n = d.numel()
    if n == 0 or top_k <= 0:
        return []
    key = d * (1 << 32) + ((1 << 32) - 1 - t.arange(n, device=d.device))
    k = min(top_k, int((d > 0).sum().item()))
    if k == 0:
        return []
    vals = t.topk(key, k).values.cpu().numpy()
    return [{"node": int((1 << 32) - 1 - (v & 0xFFFFFFFF)), "degree": int(v >> 32)}
            for v in vals]


def analyze_graph_gpu(graph_path: str, output_dir: str, dedup: bool = False,
                      top_k: int = 20) -> dict:
    """``analyze_graph_gpu`` (experiment/gpu_graph_stats.py:186-320) on the device."""
    t = _lib.require_device()
    os.makedirs(output_dir, exist_ok=True)
    t0 = time.perf_counter()
    pairs = ingest.load_edge_list_gpu(graph_path, comment_chars=COMMENT_CHARS)
    raw_edges = int(pairs.shape[0])
    if raw_edges == 0:
        raise ValueError("No edges found.")
    max_node_id = int(pairs.max().item())
    n_slots = max_node_id + 1
    t.cuda.synchronize()
    load_time = time.perf_counter() - t0

    t0 = time.perf_counter()
    if dedup:
        # unique canonical edges: the device CSR build dedups; its undirected
        # degrees are scattered back to original ids
        csr = gpu_graph.build_csr_gpu(pairs)
        n_edges = csr.n_oriented
        degree = t.zeros(n_slots, dtype=t.int32, device="cuda")
        degree[csr.node_ids] = csr.undirected_deg
        del csr
    else:
        n_edges = raw_edges
        degree = degree_count_gpu(pairs, n_slots)
    t.cuda.synchronize()
    degree_time = time.perf_counter() - t0
    del pairs

    t0 = time.perf_counter()
    summ = degree_summary_gpu(degree)
    stats_time = time.perf_counter() - t0
    t0 = time.perf_counter()
    hist, log_hist = degree_histograms_gpu(degree, summ["max_degree"])
    hist_time = time.perf_counter() - t0
    t0 = time.perf_counter()
    top_nodes = top_degrees_gpu(degree, top_k)
    topk_time = time.perf_counter() - t0

    active = summ["active_nodes"]
    degree_csv = os.path.join(output_dir, "degree_distribution.csv")
    log_csv = os.path.join(output_dir, "degree_log2_buckets.csv")
    json_path = os.path.join(output_dir, "graph_stats.json")
    write_degree_csv(degree_csv, hist, active)
    write_log_csv(log_csv, log_hist, active)
    result = {
        "graph": graph_path,
        "raw_edges": raw_edges,
        "used_edges": n_edges,
        "dedup": dedup,
        "active_nodes": active,
        "max_node_id": max_node_id,
        "degree_sum": summ["degree_sum"],
        "avg_degree": summ["avg_degree"],
        "min_degree": summ["min_degree"],
        "max_degree": summ["max_degree"],
        "std_degree": summ["std_degree"],
        "density_active_nodes": summ["density_active_nodes"],
        "degree_1_nodes": int(hist[1]) if len(hist) > 1 else 0,
        "top_degrees": top_nodes,
        # the reference's keys; the host-to-device copy happens inside the
        # load (the text is copied, then parsed on the device)
        "timing_seconds": {
            "load_cpu": load_time,
            "h2d_and_alloc": 0.0,
            "degree_kernel": degree_time,
            "stats_kernel": stats_time,
            "histogram_kernel": hist_time,
            "degree_d2h_for_topk": topk_time,
        },
        "outputs": {
            "degree_distribution_csv": degree_csv,
            "degree_log2_buckets_csv": log_csv,
            "summary_json": json_path,
        },
    }
    with open(json_path, "w") as f:
        json.dump(result, f, indent=2)
    return result
