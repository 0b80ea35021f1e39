"""CSR construction and degree orientation (drop-in for the reference's ``src/gpu_graph.py``).

``to_csr_numpy`` / ``to_csr_gpu`` keep the reference signatures
(src/gpu_graph.py:6-52).  ``build_csr_gpu`` is the device path from a raw edge
list: the whole loader of experiment/exp_gpu.py:264-389 (self-loop removal,
dedup, dense relabel, (degree, id) orientation, row-sorted CSR) runs as HIP
kernels through ``tc_csr_build_begin`` / ``tc_csr_build_finish``.
"""

from __future__ import annotations

import ctypes
from typing import Dict, NamedTuple

import numpy as np

from . import _lib


def to_csr_numpy(forward_adj: Dict[int, np.ndarray]):
    """Forward-adjacency dict -> dense int32 CSR on the host (src/gpu_graph.py:6-41).

    Returns ``(row_ptr, col_idx, degrees, node_ids, edge_u, edge_v)``, all
    int32: node ids sorted ascending, dense index = rank of the id, rows in
    the order the dict lists them (sorted in an oriented dict).
    """
    node_ids = np.array(sorted(forward_adj.keys()), dtype=np.int64)
    n = len(node_ids)
    lists = [np.asarray(forward_adj[int(x)], dtype=np.int64) for x in node_ids]
    degrees = np.array([len(a) for a in lists], dtype=np.int32)
    row_ptr = np.zeros(n + 1, dtype=np.int32)
    row_ptr[1:] = np.cumsum(degrees)
    flat = np.concatenate(lists) if n else np.zeros(0, dtype=np.int64)
    col_idx = np.searchsorted(node_ids, flat).astype(np.int32)
    if len(flat) and not np.array_equal(node_ids[np.minimum(col_idx, max(n - 1, 0))], flat):
        raise KeyError("forward_adj names a neighbour that is not a key")
    edge_u = np.repeat(np.arange(n, dtype=np.int32), degrees)
    edge_v = col_idx.copy()
    return row_ptr, col_idx, degrees, node_ids.astype(np.int32), edge_u, edge_v


def to_csr_gpu(forward_adj: Dict[int, np.ndarray]):
    """Same 6-tuple as ``to_csr_numpy`` as device tensors (src/gpu_graph.py:44-52)."""
    _lib.require_device()
    return tuple(_lib.as_device(a) for a in to_csr_numpy(forward_adj))


class DeviceCSR(NamedTuple):
    """Forward CSR on the device (fields of experiment/exp_gpu.py:388-389)."""

    row_ptr: object         # int32 or int64 [n+1]
    col_idx: object         # int32 [E], rows sorted
    degrees: object         # int32 [n] forward degree
    node_ids: object        # int64 [n] original id of each dense id
    edge_u: object          # int32 [E]
    edge_v: object          # int32 [E]
    undirected_deg: object  # int32 [n]
    n_nodes: int
    n_oriented: int


@_lib.on_stream
def build_csr_gpu(pairs, row_ptr_dtype: str = "auto", stream=None) -> DeviceCSR:
    """Raw undirected pairs (``int32[m, 2]``, host or device) -> oriented CSR on the GPU.

    ``row_ptr_dtype``: "int32", "int64" or "auto" (int64 when the edge count
    does not fit int32).
    """
    t = _lib.require_device()
    lib = _lib.load()
    p = _lib.as_device(pairs, t.int32, "pairs")
    if p.ndim != 2 or p.shape[1] != 2:
        raise ValueError(f"pairs must have shape (m, 2), got {tuple(p.shape)}")
    s = _lib.stream_ptr(stream)
    state = ctypes.c_void_p()
    n = ctypes.c_int64()
    E = ctypes.c_int64()
    _lib.check(lib.tc_csr_build_begin(_lib.ptr(p), p.shape[0], s, ctypes.byref(state),
                                      ctypes.byref(n), ctypes.byref(E)))
    try:
        n_nodes, n_edges = n.value, E.value
        use64 = row_ptr_dtype == "int64" or (row_ptr_dtype == "auto" and n_edges > 2**31 - 1)
        dev = p.device
        row_ptr = t.empty(n_nodes + 1, dtype=t.int64 if use64 else t.int32, device=dev)
        col = t.empty(max(n_edges, 0), dtype=t.int32, device=dev)
        deg = t.empty(n_nodes, dtype=t.int32, device=dev)
        nid = t.empty(n_nodes, dtype=t.int64, device=dev)
        eu = t.empty(n_edges, dtype=t.int32, device=dev)
        ev = t.empty(n_edges, dtype=t.int32, device=dev)
        udeg = t.empty(n_nodes, dtype=t.int32, device=dev)
        _lib.check(lib.tc_csr_build_finish(state, _lib.ptr(row_ptr), int(use64), _lib.ptr(col),
                                           _lib.ptr(deg), _lib.ptr(nid), _lib.ptr(eu),
                                           _lib.ptr(ev), _lib.ptr(udeg), s))
    finally:
        t.cuda.current_stream().synchronize() if stream is None else stream.synchronize()
        lib.tc_csr_build_free(state)
    return DeviceCSR(row_ptr, col, deg, nid, eu, ev, udeg, n_nodes, n_edges)


@_lib.on_stream
def degree_buckets(row_ptr, stream=None):
    """Degree-bucket sort of the vertices: ``(bucket_start int64[34], perm int32[n])``.

    Bucket k holds the vertices with ``ceil(log2(deg+ + 1)) == k`` in
    ascending id order.
    """
    t = _lib.require_device()
    rp, is64 = _lib.row_ptr_arg(row_ptr)
    n = rp.numel() - 1
    starts = t.empty(34, dtype=t.int64, device=rp.device)
    perm = t.empty(max(n, 0), dtype=t.int32, device=rp.device)
    _lib.check(_lib.load().tc_degree_buckets(_lib.ptr(rp), is64, n, _lib.ptr(starts),
                                             _lib.ptr(perm), _lib.stream_ptr(stream)))
    return starts, perm
