"""The CPU experiment's HLL variant (cpu_exp/triangle_hll_experiment.py) in the
oracle, pinned by published splitmix64 outputs, closed forms and a per-element
restatement."""

import math

import numpy as np
import pytest

from oracle import tc_oracle as O
from triangle_counting_hll_amd import graphgen


def test_seeded_splitmix64_known_answers():
    # the first outputs of a splitmix64 generator seeded with 0 (Vigna's
    # reference implementation): state advances by the golden gamma, then mixes
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    got = [O.splitmix64_seeded_int((k * O.SM64_GOLDEN) & O.MASK64) for k in range(3)]
    assert got == want


@pytest.mark.parametrize("p", [4, 10, 16])
def test_rank_from_high_bits(p):
    assert O.rho_high_int(0, p) == 65 - p
    assert O.rho_high_int((1 << 63), p) == 1           # top bit set
    assert O.rho_high_int((1 << p), p) == 64 - p       # lowest high bit only
    assert O.rho_high_int((1 << p) - 1, p) == 65 - p   # only index bits set


@pytest.mark.parametrize("p,seed", [(4, 0), (8, 42), (12, 2**63 + 5)])
def test_vectorised_registers_equal_per_element(p, seed):
    rp, col, *_ = O.csr_from_pairs(graphgen.rmat(8, 8, seed=3))
    assert np.array_equal(O.build_registers_cpu_exp(rp, col, p, seed),
                          O.build_registers_cpu_exp_py(rp, col, p, seed))


def test_estimator_branches():
    m = 1 << 16
    assert O.estimate_cpu_exp(np.zeros(m, np.uint8)) == 0.0  # LinearCounting, V = m
    r = np.full(m, 12, np.uint8)                             # raw above 2^32 / 30
    raw = O.alpha(m) * m * m / (m * 2.0 ** -12)
    assert raw > 2**32 / 30
    assert O.estimate_cpu_exp(r) == -2**32 * math.log(1 - raw / 2**32)
    r = np.full(1024, 5, np.uint8)                           # plain raw estimate
    assert O.estimate_cpu_exp(r) == O.alpha(1024) * 1024 * 1024 / (1024 * 2.0 ** -5)
    r = np.zeros(1024, np.uint8)
    r[:100] = 3                                              # LinearCounting
    assert O.estimate_cpu_exp(r) == 1024 * math.log(1024 / 924)


def test_union_is_register_max():
    rp, col, *_ = O.csr_from_pairs(graphgen.rmat(9, 16, seed=4))
    regs = O.build_registers_cpu_exp(rp, col, 10, 7)
    a, b = regs[3], regs[5]
    assert math.isclose(O.union_estimate_cpu_exp(a, a), O.estimate_cpu_exp(a), rel_tol=1e-14)
    assert math.isclose(O.union_estimate_cpu_exp(a, b),
                        O.estimate_cpu_exp(np.maximum(a, b)), rel_tol=1e-14)
    assert O.union_estimate_cpu_exp(a, np.zeros_like(a)) == O.union_estimate_cpu_exp(a, a)


def test_hybrid_threshold_extremes():
    rp, col, fdeg, nid, eu, ev, _ = O.csr_from_pairs(graphgen.rmat(9, 8, seed=6))
    regs = O.build_registers_cpu_exp(rp, col, 8, 1)
    deg = np.diff(rp)
    tot, ne, nh, es, hs, cl = O.hybrid_count_cpu_exp(rp, col, eu, ev, deg, regs, 10**9)
    assert (nh, hs, cl) == (0, 0.0, 0) and tot == O.exact_count_csr(rp, col) and ne == len(eu)
    tot, ne, nh, es, hs, cl = O.hybrid_count_cpu_exp(rp, col, eu, ev, deg, regs, -1)
    assert ne == 0 and (tot, cl) == O.pure_hll_count_cpu_exp(eu, ev, deg, regs)
