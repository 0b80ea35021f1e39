"""ctypes binding of the C oracle (test infrastructure only, see tc_oracle.c)."""

from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    """Load ``oracle/libtc_oracle.so`` (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libtc_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"C oracle not built: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        L.tco_exact_count.restype = ctypes.c_uint64
        L.tco_exact_count.argtypes = [ctypes.c_int64, P, P, ctypes.c_int64, P, P,
                                      ctypes.c_int, ctypes.c_int]
        L.tco_edge_supports.restype = None
        L.tco_edge_supports.argtypes = [P, P, ctypes.c_int64, P, P, P, ctypes.c_int]
        L.tco_hll_registers.restype = None
        L.tco_hll_registers.argtypes = [ctypes.c_int64, P, P, ctypes.c_int, P, P, ctypes.c_int]
        L.tco_max_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def exact_count(row_ptr, col_idx, edge_u=None, edge_v=None, policy=0, nthreads=0) -> int:
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    col = np.ascontiguousarray(col_idx, dtype=np.int32)
    n = len(rp) - 1
    if edge_u is None:
        return int(lib().tco_exact_count(n, _p(rp), _p(col), 0, None, None, policy, nthreads))
    eu = np.ascontiguousarray(edge_u, dtype=np.int32)
    ev = np.ascontiguousarray(edge_v, dtype=np.int32)
    return int(lib().tco_exact_count(n, _p(rp), _p(col), len(eu), _p(eu), _p(ev), policy, nthreads))


def edge_supports(row_ptr, col_idx, edge_u, edge_v, nthreads=0) -> np.ndarray:
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    col = np.ascontiguousarray(col_idx, dtype=np.int32)
    eu = np.ascontiguousarray(edge_u, dtype=np.int32)
    ev = np.ascontiguousarray(edge_v, dtype=np.int32)
    out = np.zeros(len(eu), dtype=np.int32)
    lib().tco_edge_supports(_p(rp), _p(col), len(eu), _p(eu), _p(ev), _p(out), nthreads)
    return out


def hll_registers(row_ptr, col_idx, p, slot, n_slots, nthreads=0) -> np.ndarray:
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    col = np.ascontiguousarray(col_idx, dtype=np.int32)
    sl = np.ascontiguousarray(slot, dtype=np.int64)
    regs = np.zeros((n_slots, 1 << p), dtype=np.uint8)
    lib().tco_hll_registers(len(rp) - 1, _p(rp), _p(col), p, _p(sl), _p(regs), nthreads)
    return regs


def max_threads() -> int:
    return int(lib().tco_max_threads())
