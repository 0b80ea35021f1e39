"""The prototype GPU timing experiment on MI355X (drop-in for
``experiment1_gpu/gpu_experiment_fixed.py``).

Same functions, arguments and outputs: ``load_and_orient(path)``,
``build_hll_sketches_gpu(row_ptr, col_idx, N, p)``, ``benchmark_kernel``,
``run_gpu_experiment(N, E, row_ptr, col_idx, deg_plus, edges_u, edges_v,
gammas, p, warmup, repeat, out_dir)`` writing ``gpu_timing_results.csv`` with
the reference's columns, row names and number formats (:664-697), and the
``--graph/--gammas/--p/--warmup/--repeat/--out-dir`` command line (:704-747).

The strategies keep the prototype's arithmetic (hash mode ``proto32``:
its 32-bit hash, rank window and estimator, support ``est(u) + est(v) -
est(u | v)`` raised to 0, light iff ``min(deg+u, deg+v) <= gamma``, every
vertex sketched) and run on this package's kernels:

* ``all_exact_input`` -- the exact count over the edge list in CSR order;
* ``all_exact_sorted`` -- the same over the edges sorted by cost;
* ``all_hll`` -- every edge by its HLL support;
* ``hybrid_g<gamma>`` -- light edges exact, heavy edges HLL, routed inside the
  kernels (no partition copies).  ``exact_ms`` and ``hll_ms`` are the summed
  HIP-event times of the exact and HLL stages, ``cls_ms`` the rest of the
  call (routing, compaction, the FP64 sum), as the reference derives it.

Differences: the estimator runs in FP64 instead of FP32, and the returned
dict adds the counts the prototype leaves on the device (``exact_count``,
``hll_estimate`` and each hybrid's ``estimate``).  Graph loading is the
device ingest and CSR build; the prototype's Python loader produces the same
CSR (a (degree, id) orientation is unchanged by the monotone relabel).
"""

from __future__ import annotations

import argparse
import csv
import os
import time

import numpy as np

from . import _lib, gpu_hll, gpu_hybrid, ingest

CSV_COLUMNS = ["strategy", "gamma", "time_ms", "throughput_edges_per_sec", "speedup_vs_exact",
               "light_pct", "heavy_pct", "cls_ms", "exact_ms", "hll_ms"]
HASH_MODE = "proto32"


def prototype_alpha(m: int) -> float:
    """The prototype's alpha (gpu_experiment_fixed.py:323-324)."""
    return 0.7213 / (1.0 + 1.079 / m) if m >= 128 else \
        {16: 0.673, 32: 0.697, 64: 0.709}.get(m, 0.7213)


def load_and_orient(path):
    """``load_and_orient`` (gpu_experiment_fixed.py:36-123) on the device:
    ``(N, E, row_ptr int64, col_idx, deg_plus, edges_u, edges_v)`` as CUDA tensors."""
    (row_ptr, col_idx, degrees, edge_u, edge_v, _, _, _, n_nodes, n_oriented,
     _) = ingest.load_graph_gpu(path, keep_edges_np=False, row_ptr_dtype="int64")
    return n_nodes, n_oriented, row_ptr, col_idx, degrees, edge_u, edge_v


def build_hll_sketches_gpu(d_row_ptr, d_col_idx, N, p=8):
    """``build_hll_sketches_gpu`` (gpu_experiment_fixed.py:279-292): every
    vertex sketched with the prototype's hash; returns ``(uint8[N * 2^p], m)``."""
    t = _lib.require_device()
    regs, has = gpu_hll.alloc_sketch_arrays_device(N, p)
    gpu_hll.build_sketches(d_row_ptr, d_col_idx, N, 0, p, regs, has, accumulate=False,
                           hash_mode=HASH_MODE)
    t.cuda.synchronize()
    return regs, 1 << p


def get_gpu_name(device_id=0) -> str:
    return _lib.torch().cuda.get_device_name(device_id)


def benchmark_kernel(run, warmup=3, repeat=5, label="", before=None, log=print):
    """Median and std (ms) of ``run()`` over ``repeat`` timed calls after
    ``warmup`` untimed ones, each bracketed by HIP events on the current
    stream (gpu_experiment_fixed.py:305-330).  ``before()`` runs outside the
    events (e.g. clearing an accumulator)."""
    t = _lib.torch()
    for _ in range(warmup):
        if before:
            before()
        run()
    t.cuda.synchronize()
    times = []
    for _ in range(repeat):
        if before:
            before()
        start, end = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        start.record()
        run()
        end.record()
        end.synchronize()
        times.append(start.elapsed_time(end))
    med, std = float(np.median(times)), float(np.std(times))
    if label:
        log(f"  {label:<45} {med:>9.2f} ms  (std={std:.2f})")
    return med, std, times


def write_timing_csv(csv_path, E, exact_time, sorted_time, hll_time, gammas, hybrid_results):
    """``gpu_timing_results.csv`` with the reference's rows and formats
    (gpu_experiment_fixed.py:664-697)."""
    tput_exact = E / exact_time * 1000.0
    tput_sorted = E / sorted_time * 1000.0
    tput_hll = E / hll_time * 1000.0
    sorted_sp = exact_time / sorted_time if sorted_time > 0 else 0.0
    with open(csv_path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(CSV_COLUMNS)
        w.writerow(["all_exact_input", "", f"{exact_time:.4f}", f"{tput_exact:.0f}", "1.00",
                    "100", "0", "", "", ""])
        w.writerow(["all_exact_sorted", "", f"{sorted_time:.4f}", f"{tput_sorted:.0f}",
                    f"{sorted_sp:.4f}", "100", "0", "", "", ""])
        w.writerow(["all_hll", "", f"{hll_time:.4f}", f"{tput_hll:.0f}",
                    f"{exact_time / hll_time:.4f}", "0", "100", "", "", ""])
        for gamma in gammas:
            h = hybrid_results[gamma]
            tput_h = E / h["total_time"] * 1000.0
            w.writerow([f"hybrid_g{gamma}", gamma, f"{h['total_time']:.4f}", f"{tput_h:.0f}",
                        f"{exact_time / h['total_time']:.4f}", f"{h['light_pct']:.2f}",
                        f"{h['heavy_pct']:.2f}", f"{h['cls_time']:.4f}",
                        f"{h['light_time']:.4f}", f"{h['heavy_time']:.4f}"])


def verdict(exact_time, sorted_time, hll_time, hybrid_results):
    """The GO / NO-GO checks (gpu_experiment_fixed.py:612-636):
    ``(best_gamma, [c1, c2, c3], decision)``."""
    best_gamma = min(hybrid_results, key=lambda g: hybrid_results[g]["total_time"])
    best = hybrid_results[best_gamma]["total_time"]
    best_sp = exact_time / best
    sorted_sp = exact_time / sorted_time if sorted_time > 0 else 0.0
    checks = [best_sp > 1.5, best_sp > sorted_sp, best < hll_time]
    n = sum(checks)
    decision = "GO" if n == 3 else ("MARGINAL" if n >= 2 else "NO-GO")
    return best_gamma, checks, decision


def run_gpu_experiment(N, E, row_ptr, col_idx, deg_plus, edges_u, edges_v, gammas, p=8,
                       warmup=3, repeat=5, out_dir="results_gpu", verbose=True):
    """``run_gpu_experiment`` (gpu_experiment_fixed.py:337-701) on the device."""
    t = _lib.require_device()
    log = print if verbose else (lambda *a, **k: None)
    log(f"\n{'=' * 75}\n  GPU Experiment: N={N:,}, E={E:,}, p={p}\n  Device: {get_gpu_name(0)}")
    t0 = time.perf_counter()
    rp, _ = _lib.row_ptr_arg(row_ptr)
    col = _lib.as_device(col_idx, t.int32, "col_idx")
    deg = _lib.as_device(deg_plus, t.int32, "deg_plus")
    eu = _lib.as_device(edges_u, t.int32, "edges_u")
    ev = _lib.as_device(edges_v, t.int32, "edges_v")
    t.cuda.synchronize()
    transfer_time = time.perf_counter() - t0

    t0 = time.perf_counter()
    regs, m = build_hll_sketches_gpu(rp, col, N, p)
    sketch_time = time.perf_counter() - t0
    sketch_mb = N * m / 1024 / 1024
    has = t.ones(N, dtype=t.uint8, device=rp.device)
    log(f"  Built sketches: {sketch_mb:.1f} MB ({sketch_time:.3f}s)")

    cnt = t.zeros(1, dtype=t.int64, device=rp.device)
    est = t.zeros(1, dtype=t.float64, device=rp.device)
    st = t.zeros(3, dtype=t.int64, device=rp.device)

    def clear():
        cnt.zero_()
        est.zero_()
        st.zero_()

    # B1: all exact, CSR edge order
    exact_time, _, _ = benchmark_kernel(
        lambda: gpu_hybrid.exact_count(rp, col, eu, ev, count=cnt), warmup, repeat,
        "All-exact (input order kernel)" if verbose else "", clear, log)
    exact_count = int(cnt.item())

    # B2: all exact, edges sorted by cost min(deg+u, deg+v)
    cost = t.minimum(deg[eu.long()], deg[ev.long()])
    order = t.sort(cost, stable=True).indices
    su, sv = eu[order].contiguous(), ev[order].contiguous()
    del cost, order
    sorted_time, _, _ = benchmark_kernel(
        lambda: gpu_hybrid.exact_count(rp, col, su, sv, count=cnt), warmup, repeat,
        "All-exact (sorted-by-cost kernel)" if verbose else "", clear, log)
    if int(cnt.item()) != exact_count:
        raise RuntimeError("exact count differs between edge orders")
    del su, sv

    # B3: all HLL
    hll_time, _, _ = benchmark_kernel(
        lambda: gpu_hybrid.hybrid_count(None, None, eu, ev, deg, regs, has, 0, p,
                                        _lib.TC_PART_HLL, est, st, None,
                                        hash_mode=HASH_MODE),
        warmup, repeat, "All-HLL (pure sketch kernel)" if verbose else "", clear, log)
    hll_estimate = float(est.item())

    hybrid_results = {}
    costs = t.minimum(deg[eu.long()], deg[ev.long()])
    for gamma in gammas:
        n_light = int((costs <= gamma).sum().item())
        n_heavy = E - n_light

        def run_once():
            gpu_hybrid.hybrid_count(rp, col, eu, ev, deg, regs, has, int(gamma) + 1, p,
                                    _lib.TC_PART_BOTH, est, st, None, hash_mode=HASH_MODE)

        for _ in range(warmup):
            clear()
            run_once()
        t.cuda.synchronize()
        totals, lights, heavies = [], [], []
        _lib.timing_enable(True)
        try:
            for _ in range(repeat):
                clear()
                _lib.timing_read()
                start, end = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
                start.record()
                run_once()
                end.record()
                end.synchronize()
                stages = _lib.timing_read()
                totals.append(start.elapsed_time(end))
                lights.append(stages["classify"][0] + stages["heavy"][0])
                heavies.append(stages["hll_edges"][0])
        finally:
            _lib.timing_enable(False)
        total = float(np.median(totals))
        light_time = float(np.median(lights))
        heavy_time = float(np.median(heavies))
        s = st.cpu().numpy()
        hybrid_results[gamma] = {
            "n_light": n_light, "n_heavy": n_heavy,
            "light_pct": n_light / E * 100, "heavy_pct": n_heavy / E * 100,
            "cls_time": max(total - light_time - heavy_time, 0.0),
            "light_time": light_time, "heavy_time": heavy_time, "total_time": total,
            "speedup_vs_exact": exact_time / total if total > 0 else 0.0,
            "speedup_vs_hll": hll_time / total if total > 0 else 0.0,
            "estimate": float(est.item()), "n_raised_to_zero": int(s[2]),
        }
        log(f"  Hybrid g={gamma:<5} L={n_light / E * 100:>5.1f}% H={n_heavy / E * 100:>5.1f}%  "
            f"exact={light_time:>8.2f} hll={heavy_time:>8.2f} total={total:>9.2f} ms  "
            f"vs_exact={exact_time / total:>5.2f}x")

    best_gamma, checks, decision = verdict(exact_time, sorted_time, hll_time, hybrid_results)
    log(f"  Best hybrid: g={best_gamma}; checks {checks} -> DECISION: {decision}")
    os.makedirs(out_dir, exist_ok=True)
    csv_path = os.path.join(out_dir, "gpu_timing_results.csv")
    write_timing_csv(csv_path, E, exact_time, sorted_time, hll_time, gammas, hybrid_results)
    log(f"  CSV saved to {csv_path}")
    return {
        "exact_time": exact_time, "sorted_time": sorted_time, "hll_time": hll_time,
        "hybrid_results": hybrid_results, "best_gamma": best_gamma,
        "transfer_time_s": transfer_time, "sketch_build_time_s": sketch_time,
        "sketch_mb": sketch_mb, "exact_count": exact_count, "hll_estimate": hll_estimate,
        "decision": decision,
    }


def parse_args(argv=None):
    """The prototype's command line (gpu_experiment_fixed.py:704-718)."""
    ap = argparse.ArgumentParser(description="GPU Experiment: Hybrid HLL timing verification")
    ap.add_argument("--graph", required=True, help="Path to edge-list file")
    ap.add_argument("--gammas", nargs="+", type=int, default=[32, 64, 128, 256])
    ap.add_argument("--p", type=int, default=8, help="HLL precision (default: 8)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--repeat", type=int, default=5)
    ap.add_argument("--out-dir", default="results_gpu")
    return ap.parse_args(argv)


def main(argv=None):
    args = parse_args(argv)
    N, E, row_ptr, col_idx, deg_plus, edges_u, edges_v = load_and_orient(args.graph)
    return run_gpu_experiment(N, E, row_ptr, col_idx, deg_plus, edges_u, edges_v,
                              gammas=args.gammas, p=args.p, warmup=args.warmup,
                              repeat=args.repeat, out_dir=args.out_dir)


if __name__ == "__main__":
    main()
