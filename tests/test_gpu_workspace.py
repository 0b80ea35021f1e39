"""Per-stream scratch workspaces: results stay exact across streams, sizes and
a release of the workspaces."""

import math

import numpy as np
import pytest

from oracle import cbind
from oracle import tc_oracle as O
from triangle_counting_hll_amd import _lib, gpu_graph, gpu_hybrid
from triangle_counting_hll_amd.graphgen import rmat

pytestmark = pytest.mark.gpu


def _truth(pairs):
    rp, col, *_ = O.csr_from_pairs(pairs)
    return cbind.exact_count(rp, col)


def test_large_small_large_on_one_stream(device):
    import torch

    big, small = rmat(15, 16, seed=1), rmat(9, 8, seed=2)
    cb, cs = gpu_graph.build_csr_gpu(big), gpu_graph.build_csr_gpu(small)
    tb, ts = _truth(big), _truth(small)
    counts = []
    for csr in (cb, cs, cb, cs, cb):
        c = gpu_hybrid.exact_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v)
        counts.append(c)  # enqueued back to back, read after
    torch.cuda.synchronize()
    assert [int(c.item()) for c in counts] == [tb, ts, tb, ts, tb]


def test_two_streams_interleaved(device):
    import torch

    a, b = rmat(14, 16, seed=3), rmat(13, 16, seed=4)
    ca, cb = gpu_graph.build_csr_gpu(a), gpu_graph.build_csr_gpu(b)
    ta, tb = _truth(a), _truth(b)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    out = []
    for _ in range(3):
        # the accumulator is zeroed on the stream that counts into it
        with torch.cuda.stream(s1):
            out.append(("a", gpu_hybrid.exact_count(ca.row_ptr, ca.col_idx, ca.edge_u,
                                                    ca.edge_v, stream=s1)))
        with torch.cuda.stream(s2):
            out.append(("b", gpu_hybrid.exact_count(cb.row_ptr, cb.col_idx, cb.edge_u,
                                                    cb.edge_v, stream=s2)))
    torch.cuda.synchronize()
    for tag, c in out:
        assert int(c.item()) == (ta if tag == "a" else tb)


def test_release_then_count_again(device):
    n = 300
    pairs = np.stack(np.triu_indices(n, 1), axis=1).astype(np.int32)
    csr = gpu_graph.build_csr_gpu(pairs)
    assert gpu_hybrid.gpu_exact_triangle_count(csr.row_ptr, csr.col_idx, csr.edge_u,
                                               csr.edge_v)[0] == math.comb(n, 3)
    _lib.release_workspaces()
    assert gpu_hybrid.gpu_exact_triangle_count(csr.row_ptr, csr.col_idx, csr.edge_u,
                                               csr.edge_v)[0] == math.comb(n, 3)
