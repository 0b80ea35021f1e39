"""The no-allocation exact count (tc_exact_count_noalloc) through the C ABI:
equal to tc_exact_count on two calls in a row (the second on a smaller
graph), in edge-list and CSR mode, and a graph too large for its scratch is
rejected before any device work."""

import ctypes

import numpy as np
import pytest

from oracle import cbind
from oracle import tc_oracle as O
from triangle_counting_hll_amd import _lib, gpu_graph
from triangle_counting_hll_amd.graphgen import rmat

pytestmark = pytest.mark.gpu


def _call(fn, csr, count, edges=True, params=None):
    import torch

    rp, is64 = _lib.row_ptr_arg(csr.row_ptr)
    eu = csr.edge_u if edges else None
    ev = csr.edge_v if edges else None
    rc = fn(_lib.ptr(rp), is64, _lib.ptr(csr.col_idx), csr.n_nodes, _lib.ptr(eu), _lib.ptr(ev),
            csr.n_oriented, _lib.ptr(count),
            ctypes.byref(params) if params is not None else None,
            _lib.stream_ptr(torch.cuda.current_stream()))
    return rc


def _pair(csr, edges=True, params=None):
    """(no-allocation count, allocating count) of one graph."""
    import torch

    lib = _lib.load()
    a = torch.zeros(1, dtype=torch.int64, device="cuda")
    b = torch.zeros(1, dtype=torch.int64, device="cuda")
    _lib.check(_call(lib.tc_exact_count_noalloc, csr, a, edges, params))
    _lib.check(_call(lib.tc_exact_count, csr, b, edges, params))
    torch.cuda.synchronize()
    return int(a.item()), int(b.item())


def test_noalloc_matches_allocating_entry_point(device):
    big = rmat(17, 8, seed=21)       # about 0.9M oriented edges
    small = rmat(12, 16, seed=22)    # a smaller second call
    for pairs in (big, small):
        csr = gpu_graph.build_csr_gpu(pairs)
        rp, col, *_ = O.csr_from_pairs(pairs)
        truth = cbind.exact_count(rp, col)
        got, ref = _pair(csr)
        assert got == ref == truth
        got, ref = _pair(csr, edges=False)      # the CSR's own COO
        assert got == ref == truth


def test_noalloc_with_knobs_and_repeat(device):
    csr = gpu_graph.build_csr_gpu(rmat(15, 16, seed=23))
    params = _lib.ExactParams(16, 64, 8, 4, 0, 9, 1, 1)
    first = _pair(csr, params=params)
    again = _pair(csr, params=params)
    assert first[0] == first[1] == again[0] == again[1]
    # rows cut by the edge pass (the last ids share scratch with the planner)
    trimmed = _pair(csr, params=_lib.ExactParams(16, 64, 8, 4, 0, 9, 0, 0, 1, 1))
    assert trimmed[0] == trimmed[1] == first[0]
    # items handed out by row position from per-XCD queues (their keys and
    # queue counters also come from the static scratch)
    ordered = _pair(csr, params=_lib.ExactParams(16, 64, 8, 16, 0, 10, 0, 0, 1, 1, 1))
    assert ordered[0] == ordered[1] == first[0]


def test_noalloc_rejects_graphs_beyond_its_scratch(device):
    import torch

    csr = gpu_graph.build_csr_gpu(rmat(20, 16, seed=24))   # 15.7M edges: too large
    count = torch.full((1,), 7, dtype=torch.int64, device="cuda")
    rc = _call(_lib.load().tc_exact_count_noalloc, csr, count)
    assert rc == _lib.TC_ERR_INVALID
    assert b"scratch" in _lib.load().tc_last_error()
    torch.cuda.synchronize()
    assert int(count.item()) == 7          # nothing was launched
    # the atomic plan is refused by this entry point
    rc = _call(_lib.load().tc_exact_count_noalloc, csr, count,
               params=_lib.ExactParams(0, 0, 0, 0, 1, 0, 0, 0))
    assert rc == _lib.TC_ERR_INVALID
