"""HLL sketch build and FP64 union estimation (drop-in for the reference's ``src/gpu_hll.py``).

Public surface kept from the reference:

* ``alloc_sketch_arrays(n_nodes, p)`` (src/gpu_hll.py:128-133): zeroed host
  arrays, exactly as the reference returns them.
* ``build_sketches_kernel[grid, block](row_ptr, col_idx, n_nodes, theta, p,
  sketch_regs, has_sketch)`` (src/gpu_hll.py:47-81): the Numba launch syntax is
  kept; grid and block are accepted and ignored (the HIP launcher picks its
  own shape: one workgroup per sketched vertex, registers staged in LDS).
  Host numpy arguments are copied to the device and the results copied back
  into them, as a Numba launch does; device tensors are updated in place.

Additions: ``alloc_sketch_arrays_device``, ``build_sketches`` (device tensors,
optional compact slot layout), ``union_estimate_pairs`` (host-callable form of
the device estimator ``_union_estimate_device``, src/gpu_hll.py:88-121, in
FP64 as src/hll.py:37-54), ``splitmix64`` (device hash for tests), and the
host-side constants ``_alpha`` / ``_bit_length`` for reference parity.
"""

from __future__ import annotations

import numpy as np

from . import _lib

_P_MIN, _P_MAX = 4, 16


def _check_p(p: int) -> None:
    if not (_P_MIN <= int(p) <= _P_MAX):
        raise ValueError(f"p must be in [4, 16], got {p}")


def _alpha(m: int) -> float:
    """Bias-correction constant (src/hll.py:26-34; FP64 here, FP32 in src/gpu_hll.py:20-28)."""
    if m == 16:
        return 0.673
    if m == 32:
        return 0.697
    if m == 64:
        return 0.709
    return 0.7213 / (1.0 + 1.079 / m)


def _bit_length(w: int) -> int:
    """Bits needed to represent w (src/gpu_hll.py:31-40)."""
    return int(w).bit_length()


def alloc_sketch_arrays(n_nodes: int, p: int):
    """Zeroed host arrays ``(sketch_regs uint8[n*2^p], has_sketch uint8[n])`` (src/gpu_hll.py:128-133)."""
    m = 1 << p
    return np.zeros(n_nodes * m, dtype=np.uint8), np.zeros(n_nodes, dtype=np.uint8)


def alloc_sketch_arrays_device(n_rows: int, p: int, n_nodes: int = None, device=None):
    """Zeroed device arrays (registers zero-filled on the GPU, no host copy)."""
    t = _lib.require_device()
    m = 1 << p
    n_nodes = n_rows if n_nodes is None else n_nodes
    dev = device or t.device("cuda")
    return (t.zeros(n_rows * m, dtype=t.uint8, device=dev),
            t.zeros(n_nodes, dtype=t.uint8, device=dev))


@_lib.on_stream
def build_sketches(row_ptr, col_idx, n_nodes: int, theta: int, p: int, sketch_regs,
                   has_sketch, slot=None, accumulate: bool = True, hash_mode: str = "gpu",
                   seed: int = 0, stream=None) -> None:
    """Device sketch build in place (see ``tc_hll_build`` in include/tchll.h).

    ``hash_mode="cpu_exp"`` builds the CPU experiment's registers
    (cpu_exp/triangle_hll_experiment.py:384-398: seeded splitmix64, low-p
    index, rank from the high bits) with ``seed``.
    """
    _check_p(p)
    t = _lib.require_device()
    rp, is64 = _lib.row_ptr_arg(row_ptr)
    col = _lib.as_device(col_idx, t.int32, "col_idx")
    if int(n_nodes) != rp.numel() - 1:
        raise ValueError(f"n_nodes={n_nodes} does not match row_ptr ({rp.numel() - 1} rows)")
    if sketch_regs.dtype != t.uint8 or has_sketch.dtype != t.uint8:
        raise ValueError("sketch_regs and has_sketch must be uint8")
    if not (sketch_regs.is_cuda and has_sketch.is_cuda and sketch_regs.is_contiguous()):
        raise ValueError("sketch_regs and has_sketch must be contiguous device tensors")
    m = 1 << p
    sl = None
    if slot is not None:
        sl = _lib.as_device(slot, t.int32, "slot")
        if sl.numel() != n_nodes:
            raise ValueError("slot must have one entry per node")
    elif sketch_regs.numel() < n_nodes * m:
        raise ValueError(f"sketch_regs holds {sketch_regs.numel()} bytes, need {n_nodes * m}")
    if has_sketch.numel() < n_nodes:
        raise ValueError("has_sketch is shorter than n_nodes")
    _lib.check(_lib.load().tc_hll_build(_lib.ptr(rp), is64, _lib.ptr(col), int(n_nodes),
                                        int(theta), int(p), _lib.ptr(sketch_regs), _lib.ptr(sl),
                                        _lib.ptr(has_sketch), int(bool(accumulate)),
                                        _lib.hash_mode(hash_mode), int(seed) & ((1 << 64) - 1),
                                        _lib.stream_ptr(stream)))


class _BuildSketchesLauncher:
    """``build_sketches_kernel[grid, block](...)`` with the reference argument list."""

    def __getitem__(self, config):
        def launch(row_ptr, col_idx, n_nodes, theta, p, sketch_regs, has_sketch):
            t = _lib.require_device()
            host_out = isinstance(sketch_regs, np.ndarray) or isinstance(has_sketch, np.ndarray)
            regs_d = _lib.as_device(sketch_regs, t.uint8)
            has_d = _lib.as_device(has_sketch, t.uint8)
            build_sketches(row_ptr, col_idx, n_nodes, theta, p, regs_d, has_d, accumulate=True)
            if host_out:
                t.cuda.synchronize()
                if isinstance(sketch_regs, np.ndarray):
                    sketch_regs[...] = regs_d.cpu().numpy().reshape(sketch_regs.shape)
                if isinstance(has_sketch, np.ndarray):
                    has_sketch[...] = has_d.cpu().numpy().reshape(has_sketch.shape)
            else:
                if regs_d.data_ptr() != sketch_regs.data_ptr():
                    sketch_regs.copy_(regs_d.view_as(sketch_regs))
                if has_d.data_ptr() != has_sketch.data_ptr():
                    has_sketch.copy_(has_d.view_as(has_sketch))
        return launch

    def __call__(self, *args, **kwargs):
        raise TypeError("launch with build_sketches_kernel[grid, block](...)")


build_sketches_kernel = _BuildSketchesLauncher()


@_lib.on_stream
def union_estimate_pairs(sketch_regs, us, vs, p: int, slot=None, hash_mode: str = "gpu",
                         stream=None):
    """FP64 estimates of ``|N+(u) ∪ N+(v)|`` for each pair (device float64 tensor).

    Pass ``us == vs`` for single-sketch estimates.  ``hash_mode="cpu_exp"`` uses
    the CPU experiment's estimator (cpu_exp/triangle_hll_experiment.py:356-365,
    615-635).
    """
    _check_p(p)
    t = _lib.require_device()
    regs = _lib.as_device(sketch_regs, t.uint8)
    u = _lib.as_device(us, t.int32, "us")
    v = _lib.as_device(vs, t.int32, "vs")
    if u.numel() != v.numel():
        raise ValueError("us and vs must have the same length")
    sl = _lib.as_device(slot, t.int32, "slot") if slot is not None else None
    out = t.empty(u.numel(), dtype=t.float64, device=regs.device)
    _lib.check(_lib.load().tc_union_estimate(_lib.ptr(regs), _lib.ptr(sl), int(p), _lib.ptr(u),
                                             _lib.ptr(v), u.numel(), _lib.ptr(out),
                                             _lib.hash_mode(hash_mode), _lib.stream_ptr(stream)))
    return out


@_lib.on_stream
def splitmix64(x, stream=None):
    """Device splitmix64 finaliser of a uint64/int64 tensor (src/gpu_hll.py:12-17)."""
    t = _lib.require_device()
    xd = _lib.as_device(x)
    if xd.dtype not in (t.int64, t.uint64):
        xd = xd.to(t.int64)
    out = t.empty_like(xd)
    _lib.check(_lib.load().tc_hash64(_lib.ptr(xd), _lib.ptr(out), xd.numel(),
                                     _lib.stream_ptr(stream)))
    return out
