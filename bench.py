"""Benchmark: exact triangle counting on one MI355X, with roofline and a CPU baseline.

Default workload: Chung-Lu power-law graph, n = 3,072,441, beta = 2.1,
w_max = 33,000, 117M target edges, seed 2 (about 117.5M oriented edges after
dedup), int64 row offsets.  A step is one full exact count of the graph
already resident in HBM (all pipeline kernels: classify / light merge,
re-bucketing by center, work items, LDS-hash intersection).

Prints ONE JSON line (rank 0).  ``--gpus`` above 1 is refused: this build
runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
RMAT_WORKLOADS = {"rmat14": (14, 8, 0), "rmat20": (20, 16, 5), "rmat23": (23, 16, 4)}


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def merge_model_bytes(row_ptr: np.ndarray, col: np.ndarray, eu: np.ndarray, ev: np.ndarray,
                      light_max: int):
    """Bytes of a merge intersection of every edge: sum_e 4 d+(u) + 4 d+(v) + 16.

    This is the per-edge cost of merging both rows (the reference's model of
    the exact count).  The hash-probe kernel reads far less than this (it
    streams only the shorter row), so the figure is reported as a "merge
    equivalent" rate, never as a fraction of the HBM peak.
    Returns (bytes over all edges, bytes over the edges the heavy kernel takes).
    """
    deg = np.diff(row_ptr).astype(np.int64)
    du = deg[eu]
    dv = deg[ev]
    per = 4 * (du + dv) + 16
    heavy = (du > 0) & (dv > 0) & (du + dv > light_max)
    return int(per.sum()), int(per[heavy].sum()), int(heavy.sum())


def heavy_floor_bytes(row_ptr: np.ndarray, eu: np.ndarray, ev: np.ndarray, light_max: int,
                      chunk: int, col: np.ndarray = None, trim: bool = False) -> dict:
    """Algorithmic bytes of one launch of the hash-probe kernels (k_heavy, k_heavy_big).

    Heavy edges are those with d+u, d+v > 0 and d+u + d+v > light_max.  Each
    is assigned to its center (the endpoint with the longer row; v on ties)
    and the other row is streamed against an LDS table of the center's row:
      streamed   = sum over heavy edges of 4 * (ids of the other's row streamed)
      tables     = sum over work items of 4 * d+(center), with
                   ceil(heavy edges of the center / chunk) items per center
      words      = 8 bytes per heavy edge (its packed planner word)
    B = streamed + tables + words: every byte the kernels must read once.
    The streamed ids are min(d+u, d+v); with trim (the trim_rows knob) an
    edge whose center is v streams only u's ids <= the largest id of v's
    row, and an edge left with none has no word.
    """
    deg = np.diff(row_ptr).astype(np.int64)
    du = deg[eu]
    dv = deg[ev]
    heavy = (du > 0) & (dv > 0) & (du + dv > light_max)
    hu = eu[heavy].astype(np.int64)
    hv = ev[heavy].astype(np.int64)
    du = du[heavy]
    dv = dv[heavy]
    cv = dv >= du
    center = np.where(cv, hv, hu)
    olen = np.minimum(du, dv)
    if trim:
        # ids of u's row <= last(v): a global search over (row, id) keys
        owner = np.repeat(np.arange(len(deg), dtype=np.int64), deg)
        keys = (owner << 32) | col.astype(np.int64)
        del owner
        hi = col[row_ptr[hv[cv] + 1] - 1].astype(np.int64)
        cnt = np.searchsorted(keys, (hu[cv] << 32) | hi, side="right") - row_ptr[hu[cv]]
        del keys
        olen[cv] = cnt
        keep = olen > 0
        center = center[keep]
        olen = olen[keep]
    streamed = 4 * int(olen.sum())
    per_center = np.bincount(center, minlength=len(deg)).astype(np.int64)
    items = (per_center + chunk - 1) // chunk
    tables = 4 * int((items * deg).sum())
    words = 8 * int(len(olen))
    return {"bytes": streamed + tables + words, "streamed_rows": streamed, "center_tables": tables,
            "packed_words": words, "heavy_edges": int(len(olen)), "work_items": int(items.sum()),
            "trimmed_rows": bool(trim)}


def cpu_model() -> str:
    """Host CPU model name (from /proc/cpuinfo), or the platform string."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or platform.machine()


def cpu_baseline(rp, col, budget_s=40.0, single_budget_s=8.0, runs=5):
    """Time the C oracle (the cpu_exp exact policy, OpenMP) beside the GPU run.

    Protocol: all host cores, one warm-up run, then the median of ``runs``
    (5) runs over the full graph when the six runs fit ``budget_s`` (else the
    same protocol over every k-th edge).  One thread: the same edge pass over
    every k-th oriented edge, k chosen so that one run takes about
    ``single_budget_s`` (a full single-threaded pass over the 117M-edge graph
    takes over a minute, longer than the whole GPU bench), median of 3.
    """
    from oracle import cbind

    threads = len(os.sched_getaffinity(0))
    E = len(col)
    deg = np.diff(rp)
    eu_all = np.repeat(np.arange(len(deg), dtype=np.int32), deg)

    def timed(sel, nthreads, reps):
        eu_s = None if sel is None else eu_all[sel]
        ev_s = None if sel is None else col[sel]
        c = cbind.exact_count(rp, col, eu_s, ev_s, policy=0, nthreads=nthreads)  # warm-up
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            c = cbind.exact_count(rp, col, eu_s, ev_s, policy=0, nthreads=nthreads)
            times.append(time.perf_counter() - t0)
        return float(np.median(times)), times, int(c)

    # size the work on a 1/64 slice
    probe = slice(0, E, 64)
    t0 = time.perf_counter()
    cbind.exact_count(rp, col, eu_all[probe], col[probe], policy=0, nthreads=threads)
    est_full = max((time.perf_counter() - t0) * 64, 1e-6)
    t0 = time.perf_counter()
    cbind.exact_count(rp, col, eu_all[probe], col[probe], policy=0, nthreads=1)
    est_single = max((time.perf_counter() - t0) * 64, 1e-6)

    k1 = max(1, int(np.ceil(est_single / single_budget_s)))
    sel1 = None if k1 == 1 else slice(0, E, k1)
    n1 = len(range(0, E, k1))
    dt1, t1s, _ = timed(sel1, 1, 3)
    single = {"value": n1 / dt1, "unit": "oriented edges/s", "cores": 1, "seconds": dt1,
              "runs": len(t1s), "protocol": "1 warm-up, median of 3",
              "sample": ("all oriented edges" if k1 == 1 else
                         f"every {k1}th oriented edge ({n1} of {E}); the full graph on one "
                         f"thread would take about {est_single:.0f} s")}
    k = max(1, int(np.ceil(est_full * (runs + 1) / budget_s)))
    sel = None if k == 1 else slice(0, E, k)
    n_s = len(range(0, E, k))
    dt, ts, c = timed(sel, threads, runs)
    return {"value": n_s / dt, "unit": "oriented edges/s", "cores": threads, "cpu": cpu_model(),
            "kind": "port", "seconds": dt, "runs": len(ts), "run_seconds": ts,
            "protocol": f"1 warm-up, median of {runs}",
            "count": int(c) if k == 1 else None,
            "sample": (f"all {E} oriented edges (full graph)" if k == 1 else
                       f"every {k}th oriented edge ({n_s} of {E}), u-major order"),
            "single_thread": single}


def python_exact_c1():
    """The pure-Python forward-exact count of src/exact.py:48-98 (the oracle's
    restatement: dict of sorted rows, np.intersect1d / searchsorted with the
    r = 8 rule per edge) on the R-MAT scale-14 graph (edge factor 8, seed 0),
    one thread, 1 warm-up on a tenth of the rows, then one full run."""
    from oracle import tc_oracle as O
    from triangle_counting_hll_amd.graphgen import rmat

    rp, col, *_ = O.csr_from_pairs(rmat(14, 8, seed=0))
    adj = O.forward_adj_from_csr(rp, col)
    part = dict(list(adj.items())[: max(1, len(adj) // 10)])
    O.exact_triangle_count(part)
    t0 = time.perf_counter()
    c = O.exact_triangle_count(adj)
    dt = time.perf_counter() - t0
    return {"workload": "rmat14 (scale 14, edge factor 8, seed 0)", "oriented_edges": len(col),
            "count": int(c), "seconds": dt, "value": len(col) / dt, "unit": "oriented edges/s",
            "cores": 1, "kind": "port (pure Python, src/exact.py:48-98)"}


def copy_bandwidth_gbs(gib: int = 4, reps: int = 5):
    """Measured device-to-device copy bandwidth (read + write bytes / time)."""
    import torch

    src = torch.empty((gib << 30) // 4, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    b.synchronize()
    ms = a.elapsed_time(b) / reps
    del src, dst
    return 2 * (gib << 30) / (ms * 1e-3) / 1e9


TRAFFIC_PROFILES = {"chunglu": "profiles/exact_chunglu117M_hbm_traffic.md",
                    "rmat23": "profiles/exact_rmat23_hbm_traffic.md",
                    "batch4096": "profiles/batch4096_hbm_traffic.md"}


def measured_traffic(workload: str):
    """HBM bytes per launch of k_heavy from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE scaled by a calibration copy), or None."""
    if workload not in TRAFFIC_PROFILES:
        return None, None
    path = os.path.join(ROOT, TRAFFIC_PROFILES[workload])
    if not os.path.exists(path):
        return None, None
    traffic = hit = None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch":
            traffic = float(parts[2])
        if len(parts) == 3 and parts[0] == "l2_hit_rate":
            hit = float(parts[2])
    return traffic, hit


def _timed(fn, reps):
    import torch

    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fn()
    torch.cuda.synchronize()
    return out, (time.perf_counter() - t0) * 1e3 / reps


def run_hybrid_extra(csr, truth, reps=3, params=None):
    """Fixed-threshold hybrid on the same graph (src/gpu_hybrid.py:357-431 rule).

    HLL p = 12, sketches for forward degree >= theta = gamma (the reference
    experiment passes theta = gamma, experiment/exp_gpu.py:556-557), compact
    register slots.  gamma = 1024 is the configured threshold; gamma = 128 is
    also reported because no forward degree of this graph reaches 1024.  The
    exact part runs with ``params`` (the exact count's knobs).
    """
    import torch

    from triangle_counting_hll_amd import _lib, gpu_hll, gpu_hybrid

    out = {}
    p = 12
    deg = csr.degrees
    for gamma in (1024, 128):
        sk = deg >= gamma
        n_sk = int(sk.sum().item())
        slot = torch.full_like(deg, -1)
        slot[sk] = torch.arange(n_sk, dtype=torch.int32, device=deg.device)
        regs = torch.zeros(max(n_sk, 1) << p, dtype=torch.uint8, device=deg.device)
        has = torch.zeros_like(deg, dtype=torch.uint8)

        def build():
            gpu_hll.build_sketches(csr.row_ptr, csr.col_idx, csr.n_nodes, gamma, p, regs, has,
                                   slot=slot, accumulate=False)

        _, build_ms = _timed(build, reps)

        def count():
            c, s, fb = gpu_hybrid.hybrid_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v,
                                               csr.degrees, regs, has, gamma, p, slot=slot,
                                               params=params)
            return c, s

        (c, s), ms = _timed(count, reps)
        est = float(c.item())
        st = s.cpu().numpy()
        out[f"gamma{gamma}"] = {
            "rule": "forward-degree proxy: HLL iff min(d+u, d+v) >= gamma and both sketched "
                    "(src/hybrid.py:105-112); sketches for d+ >= gamma",
            "p": p, "sketched_vertices": n_sk, "sketch_build_ms": build_ms, "count_ms": ms,
            "exact_part_knobs": dict(params) if params else "library defaults",
            "estimate": est, "rel_err": gpu_hybrid.relative_error(est, truth),
            "n_approx": int(st[0]), "n_exact": int(st[1]), "n_clamped": int(st[2]),
            "approx_coverage": int(st[0]) / max(csr.n_oriented, 1),
        }
        del regs
    out["undirected_gamma1024"] = run_undirected_proxy(csr, truth, 1024, p, reps, params)
    return out


def run_undirected_proxy(csr, truth, gamma, p, reps, params=None):
    """The hybrid with the undirected-degree proxy: HLL iff min(D(u), D(v)) >=
    gamma, D = undirected degree (d+ plus in-degree of the oriented graph), and
    both endpoints sketched; sketches (of the forward rows, as always) for
    D >= gamma; the HLL support keeps the forward degrees, d+u + d+v - union
    clamped to [0, min].  The decision is passed as the per-edge route array;
    computing it is timed with the count."""
    import torch

    from triangle_counting_hll_amd import gpu_hll, gpu_hybrid

    deg = csr.degrees
    und = deg + torch.bincount(csr.col_idx.long(), minlength=csr.n_nodes).to(torch.int32)
    sk = und >= gamma
    n_sk = int(sk.sum().item())
    slot = torch.full_like(deg, -1)
    slot[sk] = torch.arange(n_sk, dtype=torch.int32, device=deg.device)
    regs = torch.zeros(max(n_sk, 1) << p, dtype=torch.uint8, device=deg.device)
    has = sk.to(torch.uint8)
    every = torch.empty_like(has)

    def build():
        # forward rows of the undirected-proxy vertices, whatever their d+:
        # theta 0 lists every vertex, the build skips those without a slot
        gpu_hll.build_sketches(csr.row_ptr, csr.col_idx, csr.n_nodes, 0, p, regs, every,
                               slot=slot, accumulate=False)

    _, build_ms = _timed(build, reps)

    def count():
        route = ((und[csr.edge_u.long()] >= gamma) & (und[csr.edge_v.long()] >= gamma))
        c, s, fb = gpu_hybrid.hybrid_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v,
                                           deg, regs, has, 0, p, route=route.to(torch.uint8),
                                           slot=slot, params=params)
        return c, s

    (c, s), ms = _timed(count, reps)
    est = float(c.item())
    st = s.cpu().numpy()
    return {"rule": "undirected-degree proxy: HLL iff min(D(u), D(v)) >= gamma (D = d+ + d-) "
                    "and both sketched; sketches for D >= gamma; support from d+",
            "p": p, "gamma": gamma, "sketched_vertices": n_sk, "sketch_build_ms": build_ms,
            "count_ms": ms, "exact_part_knobs": dict(params) if params else "library defaults",
            "estimate": est, "rel_err": gpu_hybrid.relative_error(est, truth),
            "n_approx": int(st[0]), "n_exact": int(st[1]), "n_clamped": int(st[2]),
            "approx_coverage": int(st[0]) / max(csr.n_oriented, 1)}


def run_doulion_extra(csr, truth):
    """DOULION on the same graph (experiment/exp_gpu.py:396-450): q = 0.10 and
    0.05 over seeds 42..51, each estimate timed end to end (sampling mask,
    sampled CSR, exact count, readback); the C oracle counts one q = 0.10
    sample on the host cores."""
    import torch

    from oracle import cbind
    from triangle_counting_hll_amd import doulion, gpu_hybrid

    out = {}
    for q in doulion.DOULION_QS:
        r = doulion.doulion_forward_gpu(csr.row_ptr, csr.col_idx, q)
        out[f"q{q:g}"] = {"mean_ms": r["mean_time"] * 1e3, "mean_estimate": r["mean_estimate"],
                          "rel_err_of_mean": gpu_hybrid.relative_error(r["mean_estimate"], truth),
                          "seeds": r["n_seeds"]}
    keep = doulion.sample_mask_gpu(csr.n_oriented, 0.10, np.random.default_rng(42))
    srp, scol = doulion.edge_subgraph(csr.row_ptr, csr.col_idx, keep)
    torch.cuda.synchronize()
    rp_h = srp.cpu().numpy().astype(np.int64)
    col_h = scol.cpu().numpy()
    t0 = time.perf_counter()
    c = cbind.exact_count(rp_h, col_h, nthreads=len(os.sched_getaffinity(0)))
    out["cpu_count_of_one_q0.1_sample"] = {"seconds": time.perf_counter() - t0,
                                           "sampled_edges": int(col_h.size), "count": int(c),
                                           "threads": len(os.sched_getaffinity(0))}
    return out


def run_degree_stats_extra(workload):
    """Degree statistics of the raw edge list (experiment/gpu_graph_stats.py:89-320):
    degree count, summary and histograms on the device; numpy bincount of the
    same pairs on the host beside it."""
    import torch

    from triangle_counting_hll_amd import graph_stats

    pairs = workload_pairs(workload)
    n_slots = int(pairs.max().item()) + 1

    def run():
        deg = graph_stats.degree_count_gpu(pairs, n_slots)
        summ = graph_stats.degree_summary_gpu(deg)
        hist, log_hist = graph_stats.degree_histograms_gpu(deg, summ["max_degree"])
        return deg, summ, hist

    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    deg, summ, hist = run()
    gpu_s = time.perf_counter() - t0
    host = pairs.cpu().numpy()
    t0 = time.perf_counter()
    ref = np.bincount(host.reshape(-1), minlength=n_slots)
    cpu_s = time.perf_counter() - t0
    ref_hist = np.bincount(ref, minlength=len(hist))
    return {"pairs": int(host.shape[0]), "gpu_ms": gpu_s * 1e3,
            "numpy_bincount_ms": cpu_s * 1e3, "numpy_threads": 1,
            "max_degree": summ["max_degree"], "active_nodes": summ["active_nodes"],
            "degrees_equal_numpy": bool(np.array_equal(deg.cpu().numpy(), ref)),
            "histogram_equal_numpy": bool(np.array_equal(hist[1:], ref_hist[1:len(hist)]))}


def run_prototype_extra(csr, truth):
    """The prototype's timing table (experiment1_gpu/gpu_experiment_fixed.py:337-701)
    on the same graph: p = 8, gammas 32..256, its own hash and estimator."""
    import tempfile

    from triangle_counting_hll_amd import gpu_hybrid, prototype

    with tempfile.TemporaryDirectory() as d:
        res = prototype.run_gpu_experiment(csr.n_nodes, csr.n_oriented, csr.row_ptr, csr.col_idx,
                                           csr.degrees, csr.edge_u, csr.edge_v,
                                           [32, 64, 128, 256], p=8, warmup=2, repeat=5,
                                           out_dir=d, verbose=False)
    if res["exact_count"] != truth:
        raise RuntimeError("prototype exact count differs from the exact pipeline")
    out = {"all_exact_input_ms": res["exact_time"], "all_exact_sorted_ms": res["sorted_time"],
           "all_hll_ms": res["hll_time"],
           "all_hll_rel_err": gpu_hybrid.relative_error(res["hll_estimate"], truth),
           "sketch_build_s": res["sketch_build_time_s"], "decision": res["decision"]}
    for g, h in res["hybrid_results"].items():
        out[f"hybrid_g{g}"] = {"ms": h["total_time"], "exact_ms": h["light_time"],
                               "hll_ms": h["heavy_time"], "cls_ms": h["cls_time"],
                               "heavy_pct": h["heavy_pct"],
                               "rel_err": gpu_hybrid.relative_error(h["estimate"], truth)}
    return out


def workload_pairs(workload: str, scale_down: int = 0):
    """Raw pairs of a named workload on the device."""
    import torch

    from triangle_counting_hll_amd.graphgen import config_pairs

    if workload.startswith("rmat"):
        # device twin of the seeded host generator (same pairs, bit for bit)
        from triangle_counting_hll_amd import batch

        scale, ef, seed = RMAT_WORKLOADS[workload]
        return batch.rmat_gpu(scale - scale_down, ef, seed)
    return torch.from_numpy(config_pairs(workload, scale_down)).cuda()


def run_ingest_extra(workload, ref_csr, reps=5):
    """Edge-list text -> raw pairs -> CSR on the device (experiment/exp_gpu.py:264-389).

    The workload's raw pairs are written as a tab-separated SNAP-style text in
    HBM (``tc_format_edge_list``), then parsed back (``tc_parse_edge_list``)
    and built into the CSR.  Checks: the parsed pairs equal the input pairs
    minus self-loops, and the CSR equals the one built from the pairs.  The
    PCIe-inclusive rate copies the text from pinned host memory first.  The
    CPU figure is the oracle's restatement of the reference's Python reader
    (_iter_edges_raw, exp_gpu.py:131-143) on the first 2M lines, one thread.
    """
    import torch

    from oracle import tc_oracle as O
    from triangle_counting_hll_amd import gpu_graph, ingest

    pairs = workload_pairs(workload)
    text = ingest.format_edge_list_gpu(pairs)
    nb = text.numel()
    out = torch.empty((ingest.max_pairs(nb), 2), dtype=torch.int32, device="cuda")

    def parse():
        return ingest._parse_device(text, nb, ingest.COMMENT_CHARS, out)

    (n, bad, _), parse_ms = _timed(parse, reps)
    keep = pairs[:, 0] != pairs[:, 1]
    round_trip = bad < 0 and torch.equal(out[:n], pairs[keep])

    def text_to_csr():
        k, _, _ = ingest._parse_device(text, nb, ingest.COMMENT_CHARS, out)
        return gpu_graph.build_csr_gpu(out[:k], row_ptr_dtype="int64")

    csr, csr_ms = _timed(text_to_csr, 2)
    same_csr = (torch.equal(csr.row_ptr, ref_csr.row_ptr)
                and torch.equal(csr.col_idx, ref_csr.col_idx))
    del csr
    host = torch.empty(nb, dtype=torch.uint8, pin_memory=True)
    host.copy_(text)
    text_h = text  # reuse the device buffer as the copy target

    def h2d_parse():
        text_h.copy_(host, non_blocking=True)
        return ingest._parse_device(text_h, nb, ingest.COMMENT_CHARS, out)

    _, h2d_ms = _timed(h2d_parse, 2)
    sample = bytes(host[: min(nb, 1 << 26)].numpy())
    cut = 0
    for _ in range(2_000_000):
        nxt = sample.find(b"\n", cut)
        if nxt < 0:
            break
        cut = nxt + 1
    t0 = time.perf_counter()
    n_cpu = sum(1 for _ in O.iter_edges_raw(sample[:cut]))
    cpu_s = time.perf_counter() - t0
    lines = int(sample[:cut].count(b"\n"))
    return {
        "text_bytes": nb, "lines": int(pairs.shape[0]), "pairs": int(n),
        "parse_ms": parse_ms, "parse_gbs": nb / (parse_ms * 1e-3) / 1e9,
        "parse_lines_per_s": pairs.shape[0] / (parse_ms * 1e-3),
        "text_to_csr_ms": csr_ms, "h2d_plus_parse_ms": h2d_ms,
        "h2d_plus_parse_gbs": nb / (h2d_ms * 1e-3) / 1e9,
        "round_trip_equal": bool(round_trip), "csr_equal_to_pairs_build": bool(same_csr),
        "cpu_reference_reader": {"lines": lines, "pairs": n_cpu, "seconds": cpu_s,
                                 "lines_per_s": lines / cpu_s, "threads": 1,
                                 "sample": "first 2M lines of the same text"},
    }


def run_batch_extra(args, reps=10):
    """4,096 graphs (half ER, half R-MAT, seed 3) in one HIP-graph launch vs per-graph calls."""
    import torch

    from oracle import cbind
    from triangle_counting_hll_amd import autotune, batch, gpu_hybrid

    spec = batch.many_graphs_spec()
    t0 = time.perf_counter()
    b = batch.build_batch(spec)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    E = b.csr.n_oriented
    key = f"batch|G={b.n_graphs}|n={b.csr.n_nodes}|E={E}"
    # item_order applies to single graphs only (a batch's items are ordered by graph)
    space = {k: v for k, v in autotune.SPACE.items() if k != "item_order"}
    params, table = autotune.tune(lambda p: batch.batch_exact_count(b, p), key, reps=2,
                                  space=space)
    g = batch.BatchCountGraph(b, params)
    _, ms = _timed(g.launch, reps)
    counts = g.counts.cpu().numpy()
    graphs = batch.per_graph_csrs(b)
    plain, plain_s, plain_dev_ms = batch.plain_per_graph_counts(b, params, graphs)
    hyb, hyb_s, hyb_build_s = batch.plain_per_graph_hybrid(b, gamma=1024, p=12, graphs=graphs)
    # CPU oracle per graph on the same packed arrays
    rp = b.csr.row_ptr.cpu().numpy().astype(np.int64)
    col = b.csr.col_idx.cpu().numpy()
    eu = b.csr.edge_u.cpu().numpy()
    t0 = time.perf_counter()
    cpu_counts = np.array([cbind.exact_count(rp, col, eu[b.edge_off[i]:b.edge_off[i + 1]],
                                             col[b.edge_off[i]:b.edge_off[i + 1]])
                           for i in range(b.n_graphs)], dtype=np.int64)
    cpu_s = time.perf_counter() - t0
    g.close()
    return {
        "graphs": b.n_graphs, "n_nodes": b.csr.n_nodes, "oriented_edges": E,
        "raw_pairs": int(spec.pair_off[-1]), "build_s": build_s,
        "ms_per_launch": ms, "ms_per_graph": ms / b.n_graphs,
        "oriented_edges_per_s": E / (ms * 1e-3),
        "plain": "one gpu_exact_triangle_count call per graph on that graph's own CSR "
                 "(ids from 0), synchronised per call",
        "plain_ms_total": plain_s * 1e3, "plain_ms_per_graph": plain_s * 1e3 / b.n_graphs,
        "speedup_vs_plain": plain_s * 1e3 / ms,
        "plain_device_ms_total": plain_dev_ms,
        "plain_device_note": "sum over graphs of the device time between HIP events around "
                             "each per-graph call (its kernels and the gaps between them)",
        "speedup_vs_plain_device_time": plain_dev_ms / ms,
        "cpu_oracle_s": cpu_s, "cpu_threads": len(os.sched_getaffinity(0)),
        "plain_hybrid": {
            "rule": "per-graph gpu_hybrid_triangle_count, p = 12, gamma = theta = 1024, "
                    "compact sketches built once over the batch",
            "ms_total": hyb_s * 1e3, "sketch_build_ms": hyb_build_s * 1e3,
            "rel_err_total": gpu_hybrid.relative_error(float(hyb.sum()), int(cpu_counts.sum())),
            "max_rel_err_per_graph": float(max(
                (gpu_hybrid.relative_error(float(e), int(c))
                 for e, c in zip(hyb, cpu_counts) if c > 0), default=0.0)),
            "graphs_without_triangles_estimated_nonzero": int(np.sum((cpu_counts == 0)
                                                                     & (hyb != 0))),
            "speedup_batched_vs_plain_hybrid": hyb_s * 1e3 / ms,
        },
        "parity_batched_vs_cpu": bool(np.array_equal(counts, cpu_counts)),
        "parity_plain_vs_cpu": bool(np.array_equal(plain, cpu_counts)),
        "tuned_params": params, "autotune_evals": len(table),
        "workspace_bytes": g.workspace_bytes,
    }


def run_router_extra(reps=3, p=12, theta=256, rel_tol=0.02, budget=0.005, gates=(0, 512)):
    """Learned router on R-MAT scale 23 (seed 4), HLL p = 12, against the
    fixed-threshold hybrid on the same graph.

    Training: labels from R-MAT scale 20 (seed 5): an edge is positive iff its
    clamped HLL support is within max(1, rel_tol * exact) of its exact
    support; FP32 training with the library's gradient and Adam kernels.
    Threshold: calibrated on a held-out graph, R-MAT scale 21 (seed 6), as the
    lowest network output whose routed edges keep the sum of absolute per-edge
    errors within ``budget`` of that graph's count.  Evaluation: the network
    runs inside the exact pass (``tc_router_hybrid_count``) on the scale-23
    graph, with and without a cost gate (candidates need min(d+u, d+v) >=
    gate); timed = sketch build + the routed count.  Beside it: the degree
    rule routing the same number of edges (the candidates with the largest
    min(d+u, d+v)), which shows what the network adds at equal coverage."""
    import torch

    from triangle_counting_hll_amd import batch, gpu_graph, gpu_hybrid, router

    out = {"p": p, "theta": theta, "label_rel_tol": rel_tol, "calibration_budget": budget,
           "calibration_rule": "absolute"}
    # ---- training data from the scale-20 graph
    t0 = time.perf_counter()
    csr = gpu_graph.build_csr_gpu(batch.rmat_gpu(20, 16, seed=5))
    sk = router.build_sketches(csr, theta, p)
    cand = router.candidate_edges(csr, sk)
    feats = router.features(csr, sk, cand)
    exact, s_hll = router.edge_errors(csr, sk, cand)
    labels = router.make_labels(exact, s_hll, rel_tol=rel_tol)
    params, norm, hist = router.train(feats, labels)
    out["train"] = {"graph": "rmat20 seed 5", "edges": csr.n_oriented,
                    "candidates": int(cand.numel()),
                    "positive_frac": float(labels.mean().item()), "loss": hist,
                    "seconds": time.perf_counter() - t0}
    del csr, sk, cand, feats, exact, s_hll, labels
    torch.cuda.empty_cache()

    # ---- threshold on the held-out scale-21 graph
    csr = gpu_graph.build_csr_gpu(batch.rmat_gpu(21, 16, seed=6))
    sk = router.build_sketches(csr, theta, p)
    cand = router.candidate_edges(csr, sk)
    exact, s_hll = router.edge_errors(csr, sk, cand)
    truth21 = int(gpu_hybrid.exact_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v).item())
    prob = router.mlp_route_bf16(params, norm, router.features(csr, sk, cand), 0.5)
    thr = router.choose_threshold(prob, s_hll - exact, truth21, budget, rule="absolute")
    out["calibration"] = {"graph": "rmat21 seed 6", "edges": csr.n_oriented,
                          "candidates": int(cand.numel()), "threshold": thr}
    del csr, sk, cand, exact, s_hll, prob
    torch.cuda.empty_cache()

    # ---- evaluation graph
    csr = gpu_graph.build_csr_gpu(batch.rmat_gpu(23, 16, seed=4))
    count = torch.zeros(1, dtype=torch.int64, device="cuda")

    def exact_fn():
        count.zero_()
        gpu_hybrid.exact_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v, count)

    _, exact_ms = _timed(exact_fn, reps)
    truth = int(count.item())
    E = csr.n_oriented
    out["eval"] = {"graph": "rmat23 seed 4", "n_nodes": csr.n_nodes, "edges": E,
                   "triangles": truth, "exact_ms": exact_ms}

    def fixed(gamma):
        def fn():
            skf = router.build_sketches(csr, gamma, p)
            c, s, fb = gpu_hybrid.hybrid_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v,
                                               csr.degrees, skf.regs, skf.has, gamma, p,
                                               slot=skf.slot)
            return c, s
        (c, s), ms = _timed(fn, reps)
        est = float(c.item())
        return {"ms": ms, "rel_err": gpu_hybrid.relative_error(est, truth),
                "n_approx": int(s[0].item()), "coverage": int(s[0].item()) / E}

    out["fixed_threshold"] = {f"gamma{g}": fixed(g) for g in (64, 128, 256, 512, 768, 1024)}

    sk = router.build_sketches(csr, theta, p)
    deg = csr.degrees
    mind = torch.minimum(deg[csr.edge_u.long()], deg[csr.edge_v.long()])
    both = (sk.has[csr.edge_u.long()] == 1) & (sk.has[csr.edge_v.long()] == 1)

    def learned(gate):
        m = router.RouterModel(params, norm, thr, p, 8, theta, min_cost_degree=gate)

        def fn():
            skl = router.build_sketches(csr, max(theta, gate), p)
            return router.routed_count(m, csr, skl)
        (est, st), ms = _timed(fn, reps)
        r = {"ms": ms, "rel_err": gpu_hybrid.relative_error(est, truth),
             "n_approx": st["n_approx"], "coverage": st["n_approx"] / E,
             "network_evaluated_on": st["n_candidates"], "min_cost_degree": gate,
             "fused": True}
        # the degree rule sending as many edges to HLL: the candidates with
        # the largest min(d+u, d+v) (ties broken by edge order)
        k = st["n_approx"]
        if k:
            key = torch.where(both & (mind >= gate), mind.to(torch.int64), torch.full_like(
                mind, -1, dtype=torch.int64))
            top = torch.topk(key * E + (E - 1 - torch.arange(E, device="cuda")), k).indices
            route = torch.zeros(E, dtype=torch.uint8, device="cuda")
            route[top] = 1
            c, s, _ = gpu_hybrid.hybrid_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v,
                                              deg, sk.regs, sk.has, 0, p, route=route,
                                              slot=sk.slot)
            r["degree_rule_same_coverage_rel_err"] = gpu_hybrid.relative_error(float(c.item()),
                                                                               truth)
            del route, top, key
        return r

    out["learned"] = {f"gate{g}": learned(g) for g in gates}
    ok_fixed = {k: v for k, v in out["fixed_threshold"].items() if v["rel_err"] <= 0.01}
    ok_learned = {k: v for k, v in out["learned"].items() if v["rel_err"] <= 0.01}
    if ok_fixed:
        kf = min(ok_fixed, key=lambda k: ok_fixed[k]["ms"])
        out["fastest_fixed_within_1pct"] = dict(ok_fixed[kf], rule=kf)
    if ok_learned:
        kl = min(ok_learned, key=lambda k: ok_learned[k]["ms"])
        out["fastest_learned_within_1pct"] = dict(ok_learned[kl], policy=kl)
    if ok_fixed and ok_learned:
        out["wall_time_learned_over_fixed_within_1pct"] = (
            out["fastest_learned_within_1pct"]["ms"] / out["fastest_fixed_within_1pct"]["ms"])
    return out


def run_batch_workload(args, info):
    """The many-graph workload as the measured step: one replay of the captured
    launch that counts all 4,096 graphs (the command line's knobs); prints one
    JSON line.  The roofline is that of the hash-probe kernels inside it, timed
    with HIP events in an eager run of the same pipeline (events cannot sit
    inside a captured graph)."""
    import torch

    from oracle import cbind
    from triangle_counting_hll_amd import _lib, batch

    spec = batch.many_graphs_spec()
    b = batch.build_batch(spec)
    torch.cuda.synchronize()
    E = b.csr.n_oriented
    params = {"light_max": args.light_max, "chunk": args.chunk,
              "waves_per_cu": args.waves_per_cu, "unroll": args.unroll,
              "trim_rows": args.trim_rows, "tail_units": args.tail_units,
              "block_min_bits": args.block_min_bits, "item_order": args.item_order}
    g = batch.BatchCountGraph(b, params)
    for _ in range(args.warmup):
        g.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.launch()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / max(args.steps, 1)
    counts = g.counts.cpu().numpy()
    _lib.timing_enable(True)
    _lib.timing_read()
    eager = batch.batch_exact_count(b, params)
    stages = _lib.timing_read()
    _lib.timing_enable(False)
    rp = b.csr.row_ptr.cpu().numpy().astype(np.int64)
    col = b.csr.col_idx.cpu().numpy()
    eu = b.csr.edge_u.cpu().numpy()
    floor = heavy_floor_bytes(rp, eu, col, args.light_max, args.chunk, col=col,
                              trim=bool(args.trim_rows))
    heavy_ms, heavy_n = stages["heavy"]
    heavy_avg = heavy_ms / max(heavy_n, 1)
    parity = None
    if not args.no_cpu:
        cpu = np.array([cbind.exact_count(rp, col, eu[b.edge_off[i]:b.edge_off[i + 1]],
                                          col[b.edge_off[i]:b.edge_off[i + 1]])
                        for i in range(b.n_graphs)], dtype=np.int64)
        parity = bool(np.array_equal(counts, cpu))
    achieved = floor["bytes"] / (heavy_avg * 1e-3) / 1e9 if heavy_n else None
    traffic, l2_hit = measured_traffic("batch4096")
    line = {
        "metric": "oriented edges/sec (batched exact count, 4,096 graphs in one captured launch)",
        "value": E / (ms * 1e-3), "unit": "oriented edges/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "ms_per_graph": ms / b.n_graphs, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded generator; graphs resident in HBM before timing)",
        "config": {"workload": "batch4096 (4,096 graphs, seed 3: half ER n in [200, 50000] "
                               "p = 8/n, half R-MAT scale 8..16 edge factor 16)",
                   "n_nodes": b.csr.n_nodes, "oriented_edges": E,
                   "triangles": int(counts.sum()), **params},
        "parity_vs_cpu_oracle": parity,
        "eager_equals_captured": bool(np.array_equal(eager.cpu().numpy(), counts)),
        "roofline": {"bound": "hbm", "kernel": "k_heavy + k_heavy_big (eager run)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "traffic_source": TRAFFIC_PROFILES["batch4096"] + " (rocprofv3 --pmc "
                                       "FETCH_SIZE and WRITE_SIZE, scaled by a 1 GiB copy's "
                                       "counts)" if traffic else None,
                     "traffic_gbs": traffic / (heavy_avg * 1e-3) / 1e9
                                    if traffic and heavy_avg else None,
                     "l2_hit_rate": l2_hit,
                     "algorithmic_bytes_per_launch": floor["bytes"],
                     "algorithmic_bytes_terms": floor, "avg_launch_ms": heavy_avg},
        "stage_ms_eager": {k: v[0] / max(v[1], 1) for k, v in stages.items() if v[1]},
        "workspace_bytes": g.workspace_bytes, "device": info["name"],
    }
    g.close()
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="chunglu",
                    choices=["chunglu", "rgg21", "rmat14", "rmat20", "rmat23", "batch4096"])
    ap.add_argument("--scale-down", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--light-max", type=int, default=32)
    ap.add_argument("--chunk", type=int, default=512)
    ap.add_argument("--waves-per-cu", type=int, default=16)
    ap.add_argument("--unroll", type=int, default=16)
    ap.add_argument("--item-order", type=int, default=0, choices=(0, 1),
                    help="1: work items handed out by the position of their middle streamed "
                         "row, one queue per XCD (single graphs)")
    ap.add_argument("--block-min-bits", type=int, default=10,
                    help="centers with at least 2^(b-1) forward ids are taken by whole "
                         "workgroups (9..13)")
    ap.add_argument("--tail-units", type=int, default=0, choices=(0, 1),
                    help="1 = each streamed row's last len %% 4 ids are read in one masked "
                         "16-byte load with its last whole ids")
    ap.add_argument("--trim-rows", type=int, default=0, choices=(0, 1),
                    help="1 = the edge pass cuts each streamed row to the center row's "
                         "largest id")
    ap.add_argument("--extras", default="all",
                    help="comma-separated extra measurements to run (hybrid_fixed_threshold, "
                         "ingest_text, batch4096, learned_router, prototype_table, doulion, "
                         "degree_stats) or 'all'")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the hybrid and many-graph measurements added to the line")
    ap.add_argument("--no-autotune", action="store_true",
                    help="count with the knob flags as given instead of tuning them first")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 1 GiB device-to-device copy first (reference point for "
                         "HBM byte counters under a profiler)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path); "
              "rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from oracle import cbind
    from triangle_counting_hll_amd import _lib, gpu_graph, gpu_hybrid

    if not torch.cuda.is_available():
        print("no ROCm GPU visible", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    info = _lib.device_info()
    _log(f"device {info}")
    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        _log("calibration copy: 1 GiB read + 1 GiB written")
    if args.workload == "batch4096":
        run_batch_workload(args, info)
        return

    t0 = time.perf_counter()
    pairs_d = workload_pairs(args.workload, args.scale_down)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    _log(f"generated {pairs_d.shape[0]} raw pairs in {t_gen:.1f}s")
    t0 = time.perf_counter()
    csr = gpu_graph.build_csr_gpu(pairs_d, row_ptr_dtype="int64")
    torch.cuda.synchronize()
    t_csr = time.perf_counter() - t0
    del pairs_d
    E = csr.n_oriented
    _log(f"device CSR: n={csr.n_nodes} E={E} in {t_csr * 1e3:.1f} ms")

    params = {"light_max": args.light_max, "chunk": args.chunk,
              "waves_per_cu": args.waves_per_cu, "unroll": args.unroll,
              "trim_rows": args.trim_rows, "tail_units": args.tail_units,
              "block_min_bits": args.block_min_bits, "item_order": args.item_order}
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    tune_evals = None
    if not args.no_autotune:
        # the merge-versus-probe crossover, the work-item tile, the
        # workgroup-phase bound, the row trimming and the reading of the
        # rows' last ids chosen by the project's
        # autotuner on this graph (coordinate descent from the flags' values,
        # each candidate the median of 5 timed counts), before the timed
        # region
        from triangle_counting_hll_amd import autotune

        def run(p):
            count.zero_()
            gpu_hybrid.exact_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v, count, p)

        space = {"light_max": (24, 32, 36, 40, 44, 48), "chunk": (128, 256, 512, 1024),
                 "block_min_bits": (9, 10, 11), "trim_rows": (0, 1), "tail_units": (0, 1),
                 "item_order": (0, 1)}
        t0 = time.perf_counter()
        # a second descent starts from the row-ordered items, which pay only
        # with short work items and no workgroup phase (three knobs at once),
        # and with more edges merged in the edge pass (its gain per step is
        # below the descent's 1% margin)
        tuned, table = autotune.tune(run, f"exact|{args.workload}|n={csr.n_nodes}|E={E}",
                                     space=space, start=dict(params, block_min_bits=10,
                                                             plan="sorted"),
                                     extra_starts=[{"item_order": 1, "chunk": 256,
                                                    "block_min_bits": 11, "light_max": 44}])
        params = {k: tuned[k] for k in ("light_max", "chunk", "waves_per_cu", "unroll",
                                        "block_min_bits", "trim_rows", "tail_units",
                                        "item_order")}
        tune_evals = len(table)
        _log(f"autotuned {params} in {time.perf_counter() - t0:.1f}s ({tune_evals} evals)")
    for _ in range(args.warmup):
        count.zero_()
        gpu_hybrid.exact_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v, count, params)
    torch.cuda.synchronize()
    gpu_count = int(count.item()) if args.warmup else None

    _lib.timing_enable(True)
    _lib.timing_read()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        count.zero_()
        gpu_hybrid.exact_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v, count, params)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stages = _lib.timing_read()
    _lib.timing_enable(False)
    gpu_count = int(count.item())
    ms_per_step = (t1 - t0) * 1e3 / max(args.steps, 1)
    _log(f"exact count {gpu_count}: {ms_per_step:.3f} ms/step; stages {stages}")

    rp = csr.row_ptr.cpu().numpy().astype(np.int64)
    col = csr.col_idx.cpu().numpy()
    eu = csr.edge_u.cpu().numpy()
    B_all, B_merge_heavy, _ = merge_model_bytes(rp, col, eu, col, params["light_max"])
    floor = heavy_floor_bytes(rp, eu, col, params["light_max"], params["chunk"], col=col,
                              trim=bool(params.get("trim_rows")))
    heavy_ms, heavy_n = stages["heavy"]
    heavy_avg_ms = heavy_ms / max(heavy_n, 1)
    achieved = floor["bytes"] / (heavy_avg_ms * 1e-3) / 1e9 if heavy_n else None
    traffic, l2_hit = (measured_traffic(args.workload) if not args.scale_down
                       else (None, None))

    copy_gbs = copy_bandwidth_gbs()
    _log(f"device copy bandwidth {copy_gbs:.0f} GB/s")
    cpu = None
    parity = None
    if not args.no_cpu:
        cpu = cpu_baseline(rp, col)
        try:
            cpu["python_exact_c1"] = python_exact_c1()
        except Exception as exc:  # reported, never silently dropped
            cpu["python_exact_c1"] = {"error": f"{type(exc).__name__}: {exc}"}
        _log(f"cpu baseline {cpu}")
        if cpu.get("count") is not None:
            parity = cpu["count"] == gpu_count
        else:
            parity = cbind.exact_count(rp, col) == gpu_count

    extras = {}
    if not args.no_extras and not args.scale_down:
        for name, fn in (("hybrid_fixed_threshold",
                          lambda: run_hybrid_extra(csr, gpu_count, params=params)),
                         ("ingest_text", lambda: run_ingest_extra(args.workload, csr)),
                         ("batch4096", lambda: run_batch_extra(args)),
                         ("learned_router", lambda: run_router_extra()),
                         ("prototype_table", lambda: run_prototype_extra(csr, gpu_count)),
                         ("doulion", lambda: run_doulion_extra(csr, gpu_count)),
                         ("degree_stats", lambda: run_degree_stats_extra(args.workload))):
            if args.extras != "all" and name not in args.extras.split(","):
                continue
            try:
                t0 = time.perf_counter()
                extras[name] = fn()
                _log(f"{name}: {extras[name]} ({time.perf_counter() - t0:.1f}s)")
            except Exception as exc:  # reported, never silently dropped
                extras[name] = {"error": f"{type(exc).__name__}: {exc}"}
                _log(f"{name} failed: {exc}")

    line = {
        "metric": "oriented edges/sec (exact triangle count)",
        "value": E / (ms_per_step * 1e-3),
        "unit": "oriented edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded generator; graph resident in HBM before timing)",
        "config": {"workload": args.workload if not args.scale_down
                   else f"{args.workload}/2^{args.scale_down}",
                   "n_nodes": csr.n_nodes, "oriented_edges": E, "triangles": gpu_count,
                   "row_ptr": "int64", **params,
                   "knobs": ("autotuned on this graph before timing "
                             f"({tune_evals} candidates timed)" if tune_evals is not None
                             else "as given on the command line")},
        "parity_vs_cpu_oracle": parity,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_heavy + k_heavy_big (LDS-hash intersection)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None,
            "traffic": traffic,
            "traffic_source": f"{TRAFFIC_PROFILES[args.workload]} (rocprofv3 --pmc "
                              "FETCH_SIZE and WRITE_SIZE, scaled by a 1 GiB copy's counts)"
                              if traffic else None,
            "traffic_gbs": traffic / (heavy_avg_ms * 1e-3) / 1e9 if traffic and heavy_n else None,
            "traffic_over_algorithmic": traffic / floor["bytes"] if traffic else None,
            "l2_hit_rate": l2_hit,
            "algorithmic_bytes_per_launch": floor["bytes"],
            "algorithmic_bytes_model": "sum over heavy edges of 4 min(d+u, d+v) (streamed "
                                       "rows) + sum over work items of 4 d+(center) (table "
                                       "builds) + 8 per heavy edge (packed words)",
            "algorithmic_bytes_terms": floor,
            "avg_launch_ms": heavy_avg_ms,
            "avg_launch_source": "HIP events on the launch stream around both kernels "
                                 "(tc_timing_enable)",
            "copy_gbs_measured": copy_gbs,
            "merge_model": {
                "note": "bytes a merge of both rows would read (4 d+u + 4 d+v + 16 per "
                        "edge); an equivalent rate, not a fraction of the peak",
                "heavy_bytes_per_launch": B_merge_heavy,
                "heavy_merge_equiv_gbs": (B_merge_heavy / (heavy_avg_ms * 1e-3) / 1e9
                                          if heavy_n else None),
                "step_bytes_all_edges": B_all,
                "step_merge_equiv_gbs": B_all / (ms_per_step * 1e-3) / 1e9,
            },
        },
        "stage_ms_per_step": {k: v[0] / max(v[1], 1) for k, v in stages.items() if v[1]},
        "csr_build_ms": t_csr * 1e3,
        "cpu_baseline": cpu,
        "device": info["name"],
        "extras": extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
