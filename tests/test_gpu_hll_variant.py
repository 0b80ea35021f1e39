"""The CPU experiment's HLL variant and the pure-HLL mode on the GPU: registers
bit-exact, estimates within 1e-12 of the oracle."""

import math

import numpy as np
import pytest

from oracle import tc_oracle as O
from triangle_counting_hll_amd import gpu_graph, gpu_hll, gpu_hybrid
from triangle_counting_hll_amd.graphgen import config_pairs, rmat

pytestmark = pytest.mark.gpu


def _setup(pairs, p, seed):
    rp, col, fdeg, nid, eu, ev, _ = O.csr_from_pairs(pairs)
    csr = gpu_graph.build_csr_gpu(pairs)
    regs, has = gpu_hll.alloc_sketch_arrays_device(csr.n_nodes, p)
    gpu_hll.build_sketches(csr.row_ptr, csr.col_idx, csr.n_nodes, 0, p, regs, has,
                           hash_mode="cpu_exp", seed=seed)
    return rp, col, eu, ev, csr, regs


@pytest.mark.parametrize("p,seed", [(4, 0), (8, 42), (10, 1), (12, 2**63 + 5), (16, 77)])
def test_registers_bit_exact(device, p, seed):
    pairs = config_pairs("rmat14")
    rp, col, eu, ev, csr, regs = _setup(pairs, p, seed)
    want = O.build_registers_cpu_exp(rp, col, p, seed)
    assert np.array_equal(regs.cpu().numpy().reshape(csr.n_nodes, -1), want)


@pytest.mark.parametrize("p", [8, 12])
def test_union_estimates(device, p):
    rp, col, eu, ev, csr, regs = _setup(rmat(12, 16, seed=9), p, 3)
    R = regs.cpu().numpy().reshape(csr.n_nodes, -1)
    got = gpu_hll.union_estimate_pairs(regs, eu[:3000], ev[:3000], p,
                                       hash_mode="cpu_exp").cpu().numpy()
    want = np.array([O.union_estimate_cpu_exp(R[a], R[b]) for a, b in zip(eu[:3000], ev[:3000])])
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)
    single = gpu_hll.union_estimate_pairs(regs, eu[:200], eu[:200], p,
                                          hash_mode="cpu_exp").cpu().numpy()
    np.testing.assert_allclose(single, [O.estimate_cpu_exp(R[a]) for a in eu[:200]], rtol=1e-12)


def test_large_range_branch_uses_two_to_the_32(device):
    import torch

    m = 1 << 16
    regs = torch.full((2 * m,), 12, dtype=torch.uint8, device="cuda")
    idx = torch.zeros(1, dtype=torch.int32, device="cuda")
    got = gpu_hll.union_estimate_pairs(regs, idx, idx + 1, 16, hash_mode="cpu_exp").item()
    assert got == pytest.approx(O.estimate_cpu_exp(np.full(m, 12, np.uint8)), rel=1e-12)
    gpu = gpu_hll.union_estimate_pairs(regs, idx, idx + 1, 16).item()
    assert gpu == pytest.approx(O.estimate_from_registers(np.full(m, 12, np.uint8)), rel=1e-12)
    assert gpu != pytest.approx(got, rel=1e-6)  # the two constants differ


@pytest.mark.parametrize("mode", ["gpu", "cpu_exp"])
def test_pure_hll_count(device, mode):
    pairs = rmat(12, 16, seed=2)
    p = 10
    rp, col, fdeg, nid, eu, ev, _ = O.csr_from_pairs(pairs)
    csr = gpu_graph.build_csr_gpu(pairs)
    regs, has = gpu_hll.alloc_sketch_arrays_device(csr.n_nodes, p)
    gpu_hll.build_sketches(csr.row_ptr, csr.col_idx, csr.n_nodes, 0, p, regs, has,
                           hash_mode=mode, seed=5)
    truth = O.exact_count_csr(rp, col)
    est, secs, st = gpu_hybrid.gpu_pure_hll_triangle_count(csr.edge_u, csr.edge_v, csr.degrees,
                                                           regs, p, hash_mode=mode, truth=truth)
    deg = np.diff(rp)
    if mode == "cpu_exp":
        R = regs.cpu().numpy().reshape(csr.n_nodes, -1)
        want, clamps = O.pure_hll_count_cpu_exp(eu, ev, deg, R)
    else:
        ro, ho = O.build_registers(rp, col, p, 0)
        res = O.hybrid_count(rp, col, eu, ev, ro, ho, 0)
        want, clamps = res["total_fsum"], res["stats"]["n_clamped"]
    assert st["n_approx"] == len(eu) and st["n_clamped"] == clamps
    assert math.isclose(est, want, rel_tol=1e-12)
    assert st["rel_err"] == gpu_hybrid.relative_error(est, truth)


@pytest.mark.parametrize("threshold", [-1, 4, 16, 10**6])
def test_cpu_exp_hybrid(device, threshold):
    pairs = rmat(12, 16, seed=8)
    p = 8
    rp, col, eu, ev, csr, regs = _setup(pairs, p, 11)
    R = regs.cpu().numpy().reshape(csr.n_nodes, -1)
    got = gpu_hybrid.cpu_exp_hybrid_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v,
                                          csr.degrees, regs, p, threshold)
    want = O.hybrid_count_cpu_exp(rp, col, eu, ev, np.diff(rp), R, threshold)
    assert got[1:3] == want[1:3] and got[5] == want[5] and got[3] == want[3]
    assert math.isclose(got[4], want[4], rel_tol=1e-12, abs_tol=1e-9)
    assert math.isclose(got[0], want[0], rel_tol=1e-12, abs_tol=1e-9)
