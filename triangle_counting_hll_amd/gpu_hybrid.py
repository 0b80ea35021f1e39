"""Exact and hybrid (exact / HLL) triangle counting (drop-in for the reference's ``src/gpu_hybrid.py``).

Same public surface as src/gpu_hybrid.py:

* ``_THREADS``, ``_require_supported_block_size()``,
  ``_warn_if_float64_atomic_slow(mod)`` (src/gpu_hybrid.py:13-39);
* launch objects ``exact_count_kernel``, ``hybrid_count_kernel``,
  ``hybrid_exact_part_kernel``, ``hybrid_hll_part_kernel`` taking the
  reference argument lists with Numba ``[grid, block]`` syntax
  (src/gpu_hybrid.py:96-325).  Grid and block are accepted and ignored: the
  HIP pipeline chooses its own launch shapes.  Like the reference kernels they
  ADD into ``counts`` (and ``stats``);
* ``gpu_exact_triangle_count(row_ptr_d, col_idx_d, edge_u_d, edge_v_d)``
  -> ``(count, seconds)`` (src/gpu_hybrid.py:332-354);
* ``gpu_hybrid_triangle_count(..., gamma, p, split_kernels=True)``
  -> ``(estimate, seconds, stats)`` (src/gpu_hybrid.py:357-431).

Differences that keep results closer to the CPU oracle than the reference GPU
path: the union estimator runs in FP64 (src/hll.py:37-54) instead of FP32, and
the FP64 total is reduced in a fixed order instead of by float atomics, so
repeated runs return identical estimates.
"""

from __future__ import annotations

import ctypes
import time
from typing import Optional

import numpy as np

from . import _lib

_THREADS = 256
_FLOAT64_ATOMIC_WARNED = False


def _require_supported_block_size() -> None:
    """Kept for API parity (src/gpu_hybrid.py:19-24)."""
    if _THREADS != 256:
        raise ValueError(
            "gpu_hybrid kernels currently require _THREADS == 256 because "
            "their shared-memory reduction buffers are statically sized to 256."
        )


def _warn_if_float64_atomic_slow(cuda_module=None) -> None:
    """No-op on gfx950 (native FP64 atomics; the FP64 sum here uses none).

    Kept for API parity with src/gpu_hybrid.py:27-39.
    """
    global _FLOAT64_ATOMIC_WARNED
    _FLOAT64_ATOMIC_WARNED = True


@_lib.on_stream
def flag_indices(flags, stream=None):
    """Indices of the nonzero entries of a uint8/bool device array, ascending
    (int32 device tensor): the device form of the prototype's edge partition
    ``cp.where(light_mask)[0]`` (experiment1_gpu/gpu_experiment_fixed.py:470)."""
    t = _lib.require_device()
    f = _lib.as_device(flags, name="flags")
    if f.dtype == t.bool:
        f = f.view(t.uint8)
    if f.dtype != t.uint8:
        raise ValueError(f"flags must be uint8 or bool, got {f.dtype}")
    n = f.numel()
    idx = t.empty(max(n, 1), dtype=t.int32, device=f.device)
    cnt = t.empty(1, dtype=t.int64, device=f.device)
    _lib.check(_lib.load().tc_compact_flags(_lib.ptr(f), n, _lib.ptr(idx), _lib.ptr(cnt),
                                            _lib.stream_ptr(stream)))
    return idx[:int(cnt.item())]


def _exact_params(params: Optional[dict]):
    if not params:
        return None
    plan = params.get("plan", "sorted")
    return _lib.ExactParams(int(params.get("light_max", 0)), int(params.get("chunk", 0)),
                            int(params.get("waves_per_cu", 0)), int(params.get("unroll", 0)),
                            _lib.PLANS[plan] if isinstance(plan, str) else int(plan),
                            int(params.get("block_min_bits", 0)),
                            int(bool(params.get("pack_ids", False))),
                            int(bool(params.get("wide_offsets", False))),
                            int(bool(params.get("trim_rows", False))),
                            int(bool(params.get("tail_units", False))),
                            int(bool(params.get("item_order", False))))


@_lib.on_stream
def exact_count(row_ptr, col_idx, edge_u=None, edge_v=None, count=None, params=None,
                stream=None):
    """Enqueue the exact count; returns the device uint64 accumulator (``count`` if given)."""
    t = _lib.require_device()
    rp, is64 = _lib.row_ptr_arg(row_ptr)
    col = _lib.as_device(col_idx, t.int32, "col_idx")
    n = rp.numel() - 1
    if edge_u is None:
        # CSR mode: the edge list is the CSR's own COO, so col_idx must hold
        # exactly row_ptr[n] ids
        eu = ev = None
        E = col.numel()
        if int(rp[-1].item()) != E:
            raise ValueError(f"row_ptr[-1] = {int(rp[-1].item())} but col_idx has {E} ids")
    else:
        eu = _lib.as_device(edge_u, t.int32, "edge_u")
        ev = _lib.as_device(edge_v, t.int32, "edge_v")
        if eu.numel() != ev.numel():
            raise ValueError("edge_u and edge_v must have the same length")
        E = eu.numel()
    if count is None:
        count = t.zeros(1, dtype=t.int64, device=rp.device)
    p = _exact_params(params)
    _lib.check(_lib.load().tc_exact_count(_lib.ptr(rp), is64, _lib.ptr(col), n, _lib.ptr(eu),
                                          _lib.ptr(ev), E, _lib.ptr(count),
                                          ctypes.byref(p) if p is not None else None,
                                          _lib.stream_ptr(stream)))
    return count


@_lib.on_stream
def edge_supports(row_ptr, col_idx, edge_u, edge_v, stream=None):
    """Per-edge ``|N+(u) ∩ N+(v)|`` as a device int32 tensor."""
    t = _lib.require_device()
    rp, is64 = _lib.row_ptr_arg(row_ptr)
    col = _lib.as_device(col_idx, t.int32, "col_idx")
    eu = _lib.as_device(edge_u, t.int32, "edge_u")
    ev = _lib.as_device(edge_v, t.int32, "edge_v")
    out = t.empty(eu.numel(), dtype=t.int32, device=rp.device)
    _lib.check(_lib.load().tc_edge_supports(_lib.ptr(rp), is64, _lib.ptr(col), _lib.ptr(eu),
                                            _lib.ptr(ev), eu.numel(), _lib.ptr(out),
                                            _lib.stream_ptr(stream)))
    return out


@_lib.on_stream
def hybrid_count(row_ptr, col_idx, edge_u, edge_v, degrees, sketch_regs, has_sketch, gamma,
                 p, part=_lib.TC_PART_BOTH, counts=None, stats=None, fallback=None,
                 route=None, slot=None, per_edge=None, hash_mode: str = "gpu", params=None,
                 stream=None):
    """Enqueue one hybrid pass (``tc_hybrid_count``); returns ``(counts, stats, fallback)``.

    ``params`` (the exact count's knobs, as for ``exact_count``) selects
    ``tc_hybrid_count_tuned``: the exact part runs with those knobs.
    """
    t = _lib.require_device()
    eu = _lib.as_device(edge_u, t.int32, "edge_u")
    ev = _lib.as_device(edge_v, t.int32, "edge_v")
    if eu.numel() != ev.numel():
        raise ValueError("edge_u and edge_v must have the same length")
    dev = eu.device
    if row_ptr is not None:
        rp, is64 = _lib.row_ptr_arg(row_ptr)
        n = rp.numel() - 1
    else:
        rp, is64 = None, 0
        n = int(degrees.numel())
    col = _lib.as_device(col_idx, t.int32, "col_idx") if col_idx is not None else None
    deg = _lib.as_device(degrees, t.int32, "degrees") if degrees is not None else None
    regs = _lib.as_device(sketch_regs, t.uint8) if sketch_regs is not None else None
    has = _lib.as_device(has_sketch, t.uint8)
    rt = _lib.as_device(route, t.uint8) if route is not None else None
    sl = _lib.as_device(slot, t.int32) if slot is not None else None
    if counts is None:
        counts = t.zeros(1, dtype=t.float64, device=dev)
    if stats is None:
        stats = t.zeros(3, dtype=t.int64, device=dev)
    if fallback is None:
        fallback = t.zeros(1, dtype=t.int64, device=dev)
    args = (_lib.ptr(rp), is64, _lib.ptr(col), n, _lib.ptr(eu), _lib.ptr(ev), eu.numel(),
            _lib.ptr(deg), _lib.ptr(regs), _lib.ptr(sl), _lib.ptr(has), int(gamma), int(p),
            _lib.ptr(rt), int(part), _lib.ptr(counts), _lib.ptr(stats), _lib.ptr(fallback),
            _lib.ptr(per_edge), _lib.hash_mode(hash_mode))
    ep = _exact_params(params)
    if ep is None:
        _lib.check(_lib.load().tc_hybrid_count(*args, _lib.stream_ptr(stream)))
    else:
        _lib.check(_lib.load().tc_hybrid_count_tuned(*args, ctypes.byref(ep),
                                                      _lib.stream_ptr(stream)))
    return counts, stats, fallback


def _n(x, n_edges) -> int:
    return int(n_edges)


def _edge_slice(a, n_edges):
    t = _lib.torch()
    if isinstance(a, t.Tensor):
        return a[: int(n_edges)]
    return np.asarray(a)[: int(n_edges)]


class _Launcher:
    """Numba-style ``kernel[grid, block](*args)`` front for a HIP pipeline."""

    def __init__(self, fn, name):
        self._fn = fn
        self.__name__ = name

    def __getitem__(self, config):
        return self._fn

    def __call__(self, *args, **kwargs):
        raise TypeError(f"launch with {self.__name__}[grid, block](...)")


def _exact_kernel(row_ptr, col_idx, edge_u, edge_v, n_edges, counts):
    exact_count(row_ptr, col_idx, _edge_slice(edge_u, n_edges), _edge_slice(edge_v, n_edges),
                count=counts)


def _hybrid_fused(row_ptr, col_idx, edge_u, edge_v, n_edges, degrees, sketch_regs, has_sketch,
                  gamma, p, counts, stats):
    hybrid_count(row_ptr, col_idx, _edge_slice(edge_u, n_edges), _edge_slice(edge_v, n_edges),
                 degrees, sketch_regs, has_sketch, gamma, p, _lib.TC_PART_BOTH, counts, stats)


def _hybrid_exact_part(row_ptr, col_idx, edge_u, edge_v, n_edges, degrees, has_sketch, gamma,
                       counts, stats):
    hybrid_count(row_ptr, col_idx, _edge_slice(edge_u, n_edges), _edge_slice(edge_v, n_edges),
                 degrees, None, has_sketch, gamma, 8, _lib.TC_PART_EXACT, counts, stats)


def _hybrid_hll_part(edge_u, edge_v, n_edges, degrees, sketch_regs, has_sketch, gamma, p, counts,
                     stats):
    hybrid_count(None, None, _edge_slice(edge_u, n_edges), _edge_slice(edge_v, n_edges),
                 degrees, sketch_regs, has_sketch, gamma, p, _lib.TC_PART_HLL, counts, stats)


exact_count_kernel = _Launcher(_exact_kernel, "exact_count_kernel")
hybrid_count_kernel = _Launcher(_hybrid_fused, "hybrid_count_kernel")
hybrid_exact_part_kernel = _Launcher(_hybrid_exact_part, "hybrid_exact_part_kernel")
hybrid_hll_part_kernel = _Launcher(_hybrid_hll_part, "hybrid_hll_part_kernel")


def _sync():
    _lib.torch().cuda.synchronize()


def gpu_exact_triangle_count(row_ptr_d, col_idx_d, edge_u_d, edge_v_d, params=None):
    """Exact count and wall seconds around the device work (src/gpu_hybrid.py:332-354)."""
    _require_supported_block_size()
    t = _lib.require_device()
    rp, _ = _lib.row_ptr_arg(row_ptr_d)
    col = _lib.as_device(col_idx_d, t.int32, "col_idx")
    eu = _lib.as_device(edge_u_d, t.int32, "edge_u")
    ev = _lib.as_device(edge_v_d, t.int32, "edge_v")
    counts = t.zeros(1, dtype=t.int64, device=rp.device)
    _sync()
    t0 = time.perf_counter()
    exact_count(rp, col, eu, ev, count=counts, params=params)
    _sync()
    elapsed = time.perf_counter() - t0
    return int(counts.item()), elapsed


def gpu_hybrid_triangle_count(row_ptr_d, col_idx_d, edge_u_d, edge_v_d, degrees_d,
                              sketch_regs_d, has_sketch_d, gamma: int, p: int,
                              split_kernels: bool = True, truth: Optional[int] = None,
                              route=None, slot=None, hash_mode: str = "gpu", params=None):
    """Hybrid estimate, wall seconds and stats (src/gpu_hybrid.py:357-431).

    ``stats`` holds the reference keys ``n_approx, n_exact, n_clamped,
    approx_coverage, clamp_ratio, kernel_mode`` plus ``n_fallback`` and
    ``fallback_ratio`` (src/hybrid.py:152-164) and, when ``truth`` is given,
    ``rel_err`` (src/utils.py:36-43).  ``params`` (not in the reference):
    the exact part's knobs, as for ``gpu_exact_triangle_count``.
    """
    _require_supported_block_size()
    _warn_if_float64_atomic_slow()
    t = _lib.require_device()
    rp, _ = _lib.row_ptr_arg(row_ptr_d)
    col = _lib.as_device(col_idx_d, t.int32, "col_idx")
    eu = _lib.as_device(edge_u_d, t.int32, "edge_u")
    ev = _lib.as_device(edge_v_d, t.int32, "edge_v")
    deg = _lib.as_device(degrees_d, t.int32, "degrees")
    regs = _lib.as_device(sketch_regs_d, t.uint8)
    has = _lib.as_device(has_sketch_d, t.uint8)
    n_edges = eu.numel()
    dev = rp.device
    counts = t.zeros(1, dtype=t.float64, device=dev)
    stats_d = t.zeros(3, dtype=t.int64, device=dev)
    fb = t.zeros(1, dtype=t.int64, device=dev)
    _sync()
    t0 = time.perf_counter()
    if split_kernels:
        hybrid_count(rp, col, eu, ev, deg, None, has, gamma, p, _lib.TC_PART_EXACT, counts,
                     stats_d, fb, route=route, slot=slot, hash_mode=hash_mode, params=params)
        hybrid_count(None, None, eu, ev, deg, regs, has, gamma, p, _lib.TC_PART_HLL, counts,
                     stats_d, None, route=route, slot=slot, hash_mode=hash_mode)
    else:
        hybrid_count(rp, col, eu, ev, deg, regs, has, gamma, p, _lib.TC_PART_BOTH, counts,
                     stats_d, fb, route=route, slot=slot, hash_mode=hash_mode, params=params)
    _sync()
    elapsed = time.perf_counter() - t0
    estimate = float(counts.item())
    s = stats_d.cpu().numpy()
    n_approx, n_exact, n_clamped = int(s[0]), int(s[1]), int(s[2])
    n_fallback = int(fb.item())
    stats = {
        "n_approx": n_approx,
        "n_exact": n_exact,
        "n_clamped": n_clamped,
        "approx_coverage": n_approx / n_edges if n_edges > 0 else 0.0,
        "clamp_ratio": n_clamped / n_approx if n_approx > 0 else 0.0,
        "kernel_mode": "split_block_reduction" if split_kernels else "fused_block_reduction",
        "n_fallback": n_fallback,
        "fallback_ratio": (n_fallback / (n_approx + n_fallback)
                           if (n_approx + n_fallback) > 0 else 0.0),
    }
    if truth is not None:
        stats["rel_err"] = relative_error(estimate, truth)
    return estimate, elapsed, stats


def gpu_pure_hll_triangle_count(edge_u_d, edge_v_d, degrees_d, sketch_regs_d, p: int,
                                slot=None, hash_mode: str = "gpu",
                                truth: Optional[int] = None):
    """Every edge by its HLL support: ``(estimate, seconds, stats)``.

    The pure-HLL mode of cpu_exp/triangle_hll_experiment.py:637-660
    (``pure_hll_count``): s = deg u + deg v - union, clamped to [0, min] with
    the clamps counted; every vertex must have a register row.  ``stats`` has
    ``n_approx``, ``n_clamped``, ``clamp_ratio`` and, with ``truth``, ``rel_err``.
    """
    t = _lib.require_device()
    deg = _lib.as_device(degrees_d, t.int32, "degrees")
    eu = _lib.as_device(edge_u_d, t.int32, "edge_u")
    has = t.ones(deg.numel(), dtype=t.uint8, device=deg.device)
    counts = t.zeros(1, dtype=t.float64, device=deg.device)
    stats_d = t.zeros(3, dtype=t.int64, device=deg.device)
    _sync()
    t0 = time.perf_counter()
    hybrid_count(None, None, eu, edge_v_d, deg, sketch_regs_d, has, 0, p, _lib.TC_PART_HLL,
                 counts, stats_d, None, slot=slot, hash_mode=hash_mode)
    _sync()
    elapsed = time.perf_counter() - t0
    est = float(counts.item())
    s = stats_d.cpu().numpy()
    stats = {"n_approx": int(s[0]), "n_clamped": int(s[2]),
             "clamp_ratio": int(s[2]) / int(s[0]) if s[0] else 0.0}
    if truth is not None:
        stats["rel_err"] = relative_error(est, truth)
    return est, elapsed, stats


def cpu_exp_hybrid_count(row_ptr_d, col_idx_d, edge_u_d, edge_v_d, degrees_d, sketch_regs_d,
                         p: int, threshold: int, slot=None, hash_mode: str = "cpu_exp"):
    """The CPU experiment's hybrid (cpu_exp/triangle_hll_experiment.py:662-709):
    exact when min(deg u, deg v) <= threshold, else HLL, every vertex sketched.

    Returns ``(total, exact_edges, hll_edges, exact_sum, hll_sum, clamp_count)``.
    """
    t = _lib.require_device()
    rp, _ = _lib.row_ptr_arg(row_ptr_d)
    deg = _lib.as_device(degrees_d, t.int32, "degrees")
    has = t.ones(deg.numel(), dtype=t.uint8, device=deg.device)
    ex = t.zeros(1, dtype=t.float64, device=deg.device)
    hl = t.zeros(1, dtype=t.float64, device=deg.device)
    st = t.zeros(3, dtype=t.int64, device=deg.device)
    gamma = int(threshold) + 1  # HLL iff min > threshold
    hybrid_count(rp, col_idx_d, edge_u_d, edge_v_d, deg, None, has, gamma, p,
                 _lib.TC_PART_EXACT, ex, st, None, slot=slot, hash_mode=hash_mode)
    hybrid_count(None, None, edge_u_d, edge_v_d, deg, sketch_regs_d, has, gamma, p,
                 _lib.TC_PART_HLL, hl, st, None, slot=slot, hash_mode=hash_mode)
    exact_sum, hll_sum = float(ex.item()), float(hl.item())
    s = st.cpu().numpy()
    return (exact_sum + hll_sum, int(s[1]), int(s[0]), exact_sum, hll_sum, int(s[2]))


def relative_error(estimate: float, truth: int) -> float:
    """``|estimate - truth| / truth``; inf if truth == 0 != estimate (src/utils.py:36-43)."""
    if truth == 0:
        return float("inf") if estimate != 0 else 0.0
    return abs(estimate - truth) / truth
