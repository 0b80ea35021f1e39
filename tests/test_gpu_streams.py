"""Wrappers given ``stream=`` without the caller entering torch's stream context.

Each wrapper runs its allocations, zero fills, library calls and readbacks on
the given stream, after that stream has waited for the caller's current
stream; results read back inside a wrapper are complete, and device tensors it
returns are complete once the caller waits for the stream.
"""

import numpy as np
import pytest

from oracle import cbind
from oracle import tc_oracle as O
from triangle_counting_hll_amd import gpu_graph, gpu_hll, gpu_hybrid, router
from triangle_counting_hll_amd.graphgen import rmat

pytestmark = pytest.mark.gpu


def test_side_stream_wrappers_match_default_stream(device):
    import torch

    side = torch.cuda.Stream()
    assert side != torch.cuda.current_stream()
    pairs = rmat(14, 16, seed=7)
    rp, col, *_ = O.csr_from_pairs(pairs)
    truth = cbind.exact_count(rp, col)
    # inputs produced on the current stream, consumed on the side stream
    csr = gpu_graph.build_csr_gpu(pairs)
    cnt = gpu_hybrid.exact_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v, stream=side)
    torch.cuda.current_stream().wait_stream(side)
    assert int(cnt.item()) == truth

    # a readback inside the wrapper (the compacted length) waits for its kernels
    flags = (torch.arange(3_000_000, device="cuda") % 7 == 3).to(torch.uint8)
    idx = gpu_hybrid.flag_indices(flags, stream=side)
    torch.cuda.current_stream().wait_stream(side)
    assert np.array_equal(idx.cpu().numpy(), np.flatnonzero(flags.cpu().numpy()))

    # the hybrid: counts zeroed and accumulated on the side stream
    p, gamma = 10, 16
    sk = router.build_sketches(csr, gamma, p)
    c0, s0, _ = gpu_hybrid.hybrid_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v,
                                        csr.degrees, sk.regs, sk.has, gamma, p, slot=sk.slot)
    c1, s1, _ = gpu_hybrid.hybrid_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v,
                                        csr.degrees, sk.regs, sk.has, gamma, p, slot=sk.slot,
                                        stream=side)
    torch.cuda.current_stream().wait_stream(side)
    assert float(c1.item()) == float(c0.item())
    assert np.array_equal(s1.cpu().numpy(), s0.cpu().numpy())

    # the learned router's count reads its results back on the side stream
    model = router.RouterModel(router.init_params(1), np.array([0] * 6 + [1] * 6, np.float32),
                               threshold=0.5, p=p, theta=gamma)
    e0, st0 = router.routed_count(model, csr, sk)
    e1, st1 = router.routed_count(model, csr, sk, stream=side)
    assert e1 == e0 and st1["n_approx"] == st0["n_approx"]

    # sketches built on the side stream equal the default-stream ones
    regs = torch.zeros_like(sk.regs)
    has = torch.zeros_like(sk.has)
    gpu_hll.build_sketches(csr.row_ptr, csr.col_idx, csr.n_nodes, gamma, p, regs, has,
                           slot=sk.slot, accumulate=False, stream=side)
    torch.cuda.current_stream().wait_stream(side)
    assert torch.equal(regs, sk.regs) and torch.equal(has, sk.has)
