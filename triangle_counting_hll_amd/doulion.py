"""DOULION approximate triangle counting on the device (edge sampling + exact count).

Mirrors ``_doulion_forward`` (experiment/exp_gpu.py:396-450) and ``doulion`` /
``doulion_repeated`` (src/doulion.py:11-84): keep each forward edge with
probability ``q`` using numpy's ``default_rng(seed).random(E) < q`` -- the same
draws, generated by a HIP kernel from the generator's PCG64 state
(``tc_pcg64_below``) -- count the triangles of the sampled forward CSR with the
exact pipeline (``tc_edge_subgraph`` + ``tc_exact_count``) and scale by
``1 / q**3``.  The sample reuses the original orientation, as the reference
does, so the estimate equals the reference's for every seed.
"""

from __future__ import annotations

import ctypes
import time
from typing import Optional

import numpy as np

from . import _lib, gpu_hybrid

DOULION_QS = (0.10, 0.05)  # experiment/exp_gpu.py:91
DOULION_SEEDS = 10         # experiment/exp_gpu.py:92


def _pcg64_state(rng: np.random.Generator):
    st = rng.bit_generator.state
    if st.get("bit_generator") != "PCG64":
        raise ValueError(f"DOULION sampling needs a PCG64 generator, got {st.get('bit_generator')}")
    if st.get("has_uint32"):
        raise ValueError("the generator holds a buffered 32-bit half; draw doubles only")
    return int(st["state"]["state"]), int(st["state"]["inc"])


@_lib.on_stream
def sample_mask_gpu(n: int, q: float, rng: np.random.Generator, stream=None):
    """``rng.random(n) < q`` as a uint8 CUDA tensor; ``rng`` advances by ``n`` draws."""
    t = _lib.require_device()
    state, inc = _pcg64_state(rng)
    keep = t.empty(max(n, 0), dtype=t.uint8, device="cuda")
    m64 = (1 << 64) - 1
    _lib.check(_lib.load().tc_pcg64_below(state >> 64, state & m64, inc >> 64, inc & m64, 0,
                                          n, float(q), _lib.ptr(keep), _lib.stream_ptr(stream)))
    rng.bit_generator.advance(n)
    return keep


@_lib.on_stream
def edge_subgraph(row_ptr, col_idx, keep, stream=None):
    """Forward CSR of the kept edges (CSR order): ``(row_ptr int64[n+1], col_idx int32)``."""
    t = _lib.require_device()
    rp, is64 = _lib.row_ptr_arg(row_ptr)
    col = _lib.as_device(col_idx, t.int32, "col_idx")
    n = rp.numel() - 1
    E = col.numel()
    if keep.numel() != E:
        raise ValueError(f"keep has {keep.numel()} entries for {E} edges")
    out_rp = t.empty(n + 1, dtype=t.int64, device=rp.device)
    out_col = t.empty(max(E, 1), dtype=t.int32, device=rp.device)
    kept = ctypes.c_int64()
    _lib.check(_lib.load().tc_edge_subgraph(_lib.ptr(rp), is64, _lib.ptr(col), n, E,
                                            _lib.ptr(keep), _lib.ptr(out_rp), _lib.ptr(out_col),
                                            ctypes.byref(kept), _lib.stream_ptr(stream)))
    return out_rp, out_col[:kept.value]


def doulion_gpu(row_ptr, col_idx, q: float, rng: np.random.Generator,
                params: Optional[dict] = None):
    """One DOULION estimate (src/doulion.py:11-45 on the forward edge list):
    ``(estimate, elapsed_seconds, sampled_edges)``."""
    t = _lib.require_device()
    if not 0.0 < q <= 1.0:
        raise ValueError(f"q must be in (0, 1], got {q}")
    t.cuda.synchronize()
    t0 = time.perf_counter()
    E = int(col_idx.numel()) if hasattr(col_idx, "numel") else len(col_idx)
    keep = sample_mask_gpu(E, q, rng)
    srp, scol = edge_subgraph(row_ptr, col_idx, keep)
    if scol.numel() == 0:
        t.cuda.synchronize()
        return 0.0, time.perf_counter() - t0, 0
    cnt = gpu_hybrid.exact_count(srp, scol, None, None, None, params)
    count = int(cnt.item())
    return count / (q ** 3), time.perf_counter() - t0, int(scol.numel())


def doulion_forward_gpu(row_ptr, col_idx, q: float, n_seeds: int = DOULION_SEEDS,
                        base_seed: int = 42, params: Optional[dict] = None) -> dict:
    """``_doulion_forward`` (experiment/exp_gpu.py:396-450): seeds base_seed ..
    base_seed + n_seeds - 1, each with ``numpy.random.default_rng(seed)``."""
    estimates, times = [], []
    for seed in range(base_seed, base_seed + n_seeds):
        est, elapsed, _ = doulion_gpu(row_ptr, col_idx, q, np.random.default_rng(seed), params)
        estimates.append(est)
        times.append(elapsed)
    arr_e = np.array(estimates, dtype=np.float64)
    arr_t = np.array(times, dtype=np.float64)
    return {
        "q": q, "n_seeds": n_seeds,
        "mean_estimate": float(arr_e.mean()),
        "std_estimate": float(arr_e.std()),
        "mean_time": float(arr_t.mean()),
        "std_time": float(arr_t.std()),
        "estimates": estimates, "times": times,
    }
