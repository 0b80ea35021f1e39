"""ctypes binding of ``libtchll.so`` (the C ABI declared in ``include/tchll.h``).

The library is loaded from this package directory only.  There is no CPU
fallback: when the library is missing, or no ROCm device is visible, the
functions here raise ``RuntimeError``.
"""

from __future__ import annotations

import ctypes
import functools
import inspect
import os
import threading
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtchll.so")

TC_OK = 0
TC_ERR_INVALID = 1
TC_ERR_HIP = 2
TC_ERR_OOM = 3
TC_ERR_UNSUPPORTED = 4

TC_PART_BOTH = 0
TC_PART_EXACT = 1
TC_PART_HLL = 2

P = ctypes.c_void_p
I64 = ctypes.c_int64
I32 = ctypes.c_int

# name -> (restype, argtypes); mirrors include/tchll.h
_SIGNATURES = {
    "tc_last_error": (ctypes.c_char_p, []),
    "tc_abi_version": (I32, []),
    "tc_device_info": (I32, [P, P, P, P, ctypes.c_char_p, I32]),
    "tc_release_workspaces": (I32, []),
    "tc_timing_enable": (I32, [I32]),
    "tc_timing_read": (I32, [P, P]),
    "tc_csr_build_begin": (I32, [P, I64, P, P, P, P]),
    "tc_csr_build_finish": (I32, [P, P, I32, P, P, P, P, P, P, P]),
    "tc_csr_build_free": (None, [P]),
    "tc_parse_edge_list": (I32, [P, I64, ctypes.c_char_p, P, I64, P, P, P, P]),
    "tc_format_edge_list": (I32, [P, I64, ctypes.c_char_p, ctypes.c_char_p, P, I64, P, P]),
    "tc_degree_buckets": (I32, [P, I32, I64, P, P, P]),
    "tc_exact_count": (I32, [P, I32, P, I64, P, P, I64, P, P, P]),
    "tc_exact_count_noalloc": (I32, [P, I32, P, I64, P, P, I64, P, P, P]),
    "tc_edge_supports": (I32, [P, I32, P, P, P, I64, P, P]),
    "tc_hash64": (I32, [P, P, I64, P]),
    "tc_compact_flags": (I32, [P, I64, P, P, P]),
    "tc_hll_build": (I32, [P, I32, P, I64, I64, I32, P, P, P, I32, I32, ctypes.c_uint64, P]),
    "tc_union_estimate": (I32, [P, P, I32, P, P, I64, P, I32, P]),
    "tc_hybrid_count": (I32, [P, I32, P, I64, P, P, I64, P, P, P, P, I64, I32, P, I32, P, P, P, P,
                              I32, P]),
    "tc_hybrid_count_tuned": (I32, [P, I32, P, I64, P, P, I64, P, P, P, P, I64, I32, P, I32, P, P,
                                    P, P, I32, P, P]),
    "tc_router_features": (I32, [P, P, P, P, I64, P, P, I32, I32, P, P, P]),
    "tc_router_n_params": (I32, []),
    "tc_mlp_forward": (I32, [P, P, P, I64, P, P]),
    "tc_mlp_grad": (I32, [P, P, P, P, I64, ctypes.c_float, P, P, P]),
    "tc_mlp_adam": (I32, [P, P, P, P, I32, I32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                          ctypes.c_float, P]),
    "tc_mlp_route_bf16": (I32, [P, P, P, I64, ctypes.c_float, P, P, P, P]),
    "tc_router_hybrid_count": (I32, [P, I32, P, I64, P, P, I64, P, P, P, P, I32, P, P, P, P, P,
                                     P, P]),
    "tc_vertex_graph_ids": (I32, [P, I64, P, I32, P, P]),
    "tc_batch_exact_count": (I32, [P, I32, P, I64, P, P, I64, P, I32, P, P, P]),
    "tc_batch_graph_create": (I32, [P, I32, P, I64, P, P, I64, P, I32, P, P, P, P, P]),
    "tc_batch_graph_launch": (I32, [P, P]),
    "tc_batch_graph_free": (None, [P]),
    "tc_pcg64_below": (I32, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                             ctypes.c_uint64, I64, ctypes.c_double, P, P]),
    "tc_edge_subgraph": (I32, [P, I32, P, I64, I64, P, P, P, P, P]),
    "tc_degree_count": (I32, [P, P, I64, I64, P, I64, P]),
    "tc_degree_summary": (I32, [P, I64, P, P, P]),
    "tc_degree_histogram": (I32, [P, I64, P, I64, P, I32, P]),
    "tc_generate_pairs": (I32, [P, P, P, P, P, I32, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_uint32, P, I64, P]),
}

EXPORTED = tuple(_SIGNATURES)


class ExactParams(ctypes.Structure):
    """``tc_exact_params`` of include/tchll.h (0 = library default)."""

    _fields_ = [("light_max", ctypes.c_int), ("chunk", ctypes.c_int),
                ("waves_per_cu", ctypes.c_int), ("unroll", ctypes.c_int),
                ("plan", ctypes.c_int), ("block_min_bits", ctypes.c_int),
                ("pack_ids", ctypes.c_int), ("wide_offsets", ctypes.c_int),
                ("trim_rows", ctypes.c_int), ("tail_units", ctypes.c_int),
                ("item_order", ctypes.c_int)]


class RouterNet(ctypes.Structure):
    """``tc_router_net`` of include/tchll.h (device pointers)."""

    _fields_ = [("params", ctypes.c_void_p), ("norm", ctypes.c_void_p),
                ("card", ctypes.c_void_p), ("threshold", ctypes.c_float),
                ("sample_bits", ctypes.c_int), ("min_cost_degree", ctypes.c_int64)]


PLANS = {"sorted": 0, "atomic": 1}

# hash / estimator variants of the HLL entry points (TC_HLL_* in include/tchll.h)
HASH_MODES = {"gpu": 0, "cpu_exp": 1, "proto32": 2}


def hash_mode(name) -> int:
    if name not in HASH_MODES:
        raise ValueError(f"hash_mode must be one of {sorted(HASH_MODES)}, got {name!r}")
    return HASH_MODES[name]


_lock = threading.Lock()
_lib: Optional[ctypes.CDLL] = None


def load() -> ctypes.CDLL:
    """Load the HIP library (no device work).  Raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with "
                    "`python -m triangle_counting_hll_amd.build` (there is no CPU fallback)")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in _SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(rc: int) -> None:
    """Map a TC_* status to the Python exception the reference would raise."""
    if rc == TC_OK:
        return
    msg = (load().tc_last_error() or b"").decode(errors="replace")
    if rc == TC_ERR_INVALID:
        raise ValueError(msg)
    if rc == TC_ERR_OOM:
        raise MemoryError(msg)
    raise RuntimeError(msg or f"tchll error {rc}")


# ---------------------------------------------------------------------------
# torch plumbing (device memory and streams only)
# ---------------------------------------------------------------------------

def torch():
    import torch as _t
    return _t


def require_device():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("no ROCm GPU is visible: the MI355X triangle-counting path needs one")
    return t


def stream_ptr(stream=None) -> int:
    t = torch()
    s = stream if stream is not None else t.cuda.current_stream()
    return int(s.cuda_stream)


def on_stream(fn):
    """Run a wrapper's whole body on its ``stream`` argument.

    A wrapper that takes ``stream=`` allocates and zeroes accumulators,
    converts inputs and reads results back through torch, which works on
    torch's current stream.  When ``stream`` is given and is not the current
    stream, the body runs with ``stream`` made current (so those steps and the
    library's kernels are ordered on one stream and a readback waits for the
    kernels), after ``stream`` has waited for the work already queued on the
    caller's current stream (which produced the inputs).  Device tensors the
    wrapper returns belong to ``stream``: a caller that uses them on another
    stream synchronises with ``stream`` first.
    """
    sig = inspect.signature(fn)

    @functools.wraps(fn)
    def wrapper(*args, **kwargs):
        bound = sig.bind_partial(*args, **kwargs)
        s = bound.arguments.get("stream")
        if s is None:
            return fn(*args, **kwargs)
        t = torch()
        cur = t.cuda.current_stream()
        if s == cur:
            return fn(*args, **kwargs)
        s.wait_stream(cur)
        with t.cuda.stream(s):
            return fn(*args, **kwargs)

    return wrapper


def ptr(x) -> Optional[int]:
    """Device pointer of a torch tensor (None -> NULL)."""
    if x is None:
        return None
    return int(x.data_ptr())


def as_device(x, dtype=None, name: str = "array"):
    """Return ``x`` as a contiguous CUDA tensor (copying host arrays to the GPU)."""
    t = require_device()
    if isinstance(x, np.ndarray):
        x = t.from_numpy(np.ascontiguousarray(x))
    if not isinstance(x, t.Tensor):
        x = t.as_tensor(x)
    if dtype is not None and x.dtype != dtype:
        if dtype in (t.int32,) and x.numel() and x.dtype in (t.int64,):
            if int(x.max()) > np.iinfo(np.int32).max or int(x.min()) < np.iinfo(np.int32).min:
                raise ValueError(f"{name} values do not fit int32")
        x = x.to(dtype)
    if not x.is_cuda:
        x = x.cuda()
    return x.contiguous()


def row_ptr_arg(row_ptr):
    """(tensor, is64) for a CSR row pointer of int32 or int64."""
    t = torch()
    rp = as_device(row_ptr, name="row_ptr")
    if rp.dtype == t.int64:
        return rp, 1
    if rp.dtype != t.int32:
        rp = rp.to(t.int64)
        return rp, 1
    return rp, 0


def release_workspaces() -> None:
    """Free the per-stream scratch workspaces (synchronises the device)."""
    check(load().tc_release_workspaces())


STAGES = ("classify", "heavy", "hll_build", "hll_edges")


def timing_enable(on: bool = True) -> None:
    """Record HIP events around the timed kernel stages (see tc_timing_enable)."""
    check(load().tc_timing_enable(int(bool(on))))


def timing_read() -> dict:
    """Synchronise and return ``{stage: (total_ms, launches)}``, clearing the log."""
    ms = (ctypes.c_double * len(STAGES))()
    n = (ctypes.c_int64 * len(STAGES))()
    check(load().tc_timing_read(ms, n))
    return {s: (ms[i], n[i]) for i, s in enumerate(STAGES)}


def device_info() -> dict:
    require_device()
    lib = load()
    n_cu = ctypes.c_int()
    ws = ctypes.c_int()
    lds = ctypes.c_int64()
    mem = ctypes.c_int64()
    name = ctypes.create_string_buffer(256)
    check(lib.tc_device_info(ctypes.byref(n_cu), ctypes.byref(ws), ctypes.byref(lds),
                             ctypes.byref(mem), name, 256))
    return {"n_cu": n_cu.value, "wave_size": ws.value, "lds_per_block": lds.value,
            "total_mem": mem.value, "name": name.value.decode(errors="replace")}
