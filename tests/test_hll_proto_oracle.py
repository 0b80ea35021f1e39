"""The prototype GPU experiment's HLL variant (experiment1_gpu/
gpu_experiment_fixed.py) in the oracle, pinned by the mixer's published
inverse, the reference's two rank loops, the already-pinned estimator of
src/hll.py, closed forms and the exact count."""

import math

import numpy as np
import pytest

from oracle import tc_oracle as O
from triangle_counting_hll_amd import graphgen


def _unhash32(h: int) -> int:
    # the published inverse of the 0x45D9F3B xorshift-multiply mixer:
    # multiplier 0x119DE1F3, same shifts; a wrong constant or shift in the
    # oracle's mixer breaks the round trip
    h = ((h >> 16) ^ h) * 0x119DE1F3 & O.MASK32
    h = ((h >> 16) ^ h) * 0x119DE1F3 & O.MASK32
    return (h >> 16) ^ h


def test_mixer_inverts_with_published_unhash():
    rng = np.random.default_rng(0)
    ids = [0, 1, 2, 0x7FFFFFFF, 0x9747B28C, 0xFFFFFFFF] + rng.integers(0, 2**32, 500).tolist()
    for v in ids:
        assert _unhash32(O.proto_hash32_int(v)) ^ O.PROTO_SEED32 == v
    assert (0x45D9F3B * 0x119DE1F3) % 2**32 == 1
    assert O.proto_hash32_int(O.PROTO_SEED32) == 0  # seed ^ id == 0 maps to 0


def test_vectorised_mixer_matches_int():
    v = np.random.default_rng(1).integers(0, 2**31, 5000)
    h = O.proto_hash32(v)
    assert [int(x) for x in h] == [O.proto_hash32_int(int(x)) for x in v]


def _rank_older_script(h: int, p: int) -> int:
    # experiment1_gpu/gpu_experiment.py:160-168 counts the same zeros with an
    # ascending loop over the window
    temp = ((h << p) | 1) & O.MASK32
    rho = 0
    for bit in range(32 - p):
        if temp & (1 << (31 - p - bit)):
            break
        rho += 1
    return rho + 1


@pytest.mark.parametrize("p", list(range(4, 17)))
def test_rank_window(p):
    rng = np.random.default_rng(p)
    hs = [0, 1, O.MASK32, 1 << (31 - p), 1 << (32 - 2 * p) if 32 - 2 * p >= 0 else 1,
          (1 << (32 - p)) - 1] + rng.integers(0, 2**32, 2000).tolist()
    hs += [int(x) >> s for x, s in zip(rng.integers(0, 2**32, 500), rng.integers(0, 32, 500))]
    j, r = O.proto_index_rank(np.array(hs, dtype=np.uint64), p)
    for i, h in enumerate(hs):
        assert r[i] == O.proto_rank_loop(h, p) == _rank_older_script(h, p), (h, p)
        assert j[i] == (h >> (32 - p)) & ((1 << p) - 1)
    assert r.min() >= 1 and r.max() <= 32 - p
    # the window holds the low 32 - 2p bits of h, then p - 1 zeros, then the 1
    assert O.proto_rank_loop(0, p) == 32 - p


@pytest.mark.parametrize("p", [4, 8, 12])
def test_registers_vectorised_vs_per_element(p):
    pairs = graphgen.rmat(9, 8, seed=3)
    rp, col, *_ = O.csr_from_pairs(pairs)
    assert np.array_equal(O.build_registers_proto(rp, col, p), O.build_registers_proto_py(rp, col, p))


def test_estimator_is_src_hll_without_large_range():
    # below 2^64 / 30 the large-range step of src/hll.py never fires, so the
    # prototype's estimator equals the pinned estimate_rows bit for bit
    rng = np.random.default_rng(4)
    for m in (16, 32, 64, 256, 4096):
        regs = rng.integers(0, 20, size=(50, m)).astype(np.uint8)
        regs[::3, : m // 2] = 0  # LinearCounting rows
        assert np.array_equal(O.estimate_proto_rows(regs), O.estimate_rows(regs))
    assert O.estimate_proto_rows(np.zeros((1, 256), np.uint8))[0] == 0.0


def test_support_special_cases():
    rng = np.random.default_rng(5)
    a = rng.integers(0, 12, size=(20, 256)).astype(np.uint8)
    s, neg = O.support_proto_rows(a, a)  # A with itself: est(A)
    np.testing.assert_array_equal(s, O.estimate_proto_rows(a))
    assert not neg.any()
    s, neg = O.support_proto_rows(a, np.zeros_like(a))  # A with the empty set: 0
    np.testing.assert_array_equal(s, 0.0)
    # independent rows push est(u | v) above est(u) + est(v) (raised to 0 and
    # flagged); rows that mostly repeat the other's registers keep a positive
    # support
    b = rng.integers(0, 12, size=(2000, 256)).astype(np.uint8)
    c = rng.integers(0, 12, size=(2000, 256)).astype(np.uint8)
    c[::2] = np.minimum(b[::2], c[::2])
    s, neg = O.support_proto_rows(b, c)
    raw = (O.estimate_proto_rows(b) + O.estimate_proto_rows(c)
           - O.estimate_proto_rows(np.maximum(b, c)))
    assert neg.any() and (~neg).any()
    np.testing.assert_array_equal(neg, raw < 0)
    np.testing.assert_array_equal(s, np.where(raw < 0, 0.0, raw))  # no upper clamp


def test_fp64_restatement_within_fp32_rounding_of_the_kernel():
    pairs = graphgen.rmat(11, 16, seed=6)
    rp, col, fdeg, nid, eu, ev, _ = O.csr_from_pairs(pairs)
    for p in (4, 8):
        R = O.build_registers_proto(rp, col, p)
        s64, _ = O.support_proto_rows(R[eu[:300]], R[ev[:300]])
        scale = np.maximum(O.estimate_proto_rows(R[eu[:300]]), O.estimate_proto_rows(R[ev[:300]]))
        s32 = np.array([O.support_proto_fp32(R[a], R[b]) for a, b in zip(eu[:300], ev[:300])])
        assert np.all(np.abs(s32 - s64) <= 1e-5 * scale + 1e-6)


def test_hybrid_limits_and_split():
    pairs = graphgen.rmat(11, 16, seed=7)
    rp, col, fdeg, nid, eu, ev, _ = O.csr_from_pairs(pairs)
    R = O.build_registers_proto(rp, col, 8)
    truth = O.exact_count_csr(rp, col)
    deg = np.diff(rp)
    all_exact = O.hybrid_count_proto(rp, col, eu, ev, R, int(deg.max()))
    assert all_exact["n_heavy"] == 0 and all_exact["total_fsum"] == truth
    all_hll = O.hybrid_count_proto(rp, col, eu, ev, R, -1)
    assert all_hll["n_light"] == 0
    s, neg = O.support_proto_rows(R[eu], R[ev])
    assert math.isclose(all_hll["total_fsum"], math.fsum(s.tolist()), rel_tol=1e-15)
    assert all_hll["n_raised"] == int(neg.sum())
    assert all_hll["n_raised_min"] <= all_hll["n_raised"] <= all_hll["n_raised_max"]
    # exact-zero ties exist in LinearCounting (m * V(u|v) == V(u) * V(v))
    assert all_hll["n_raised_max"] - all_hll["n_raised_min"] < 0.01 * len(eu)
    mid = O.hybrid_count_proto(rp, col, eu, ev, R, 8)
    light = np.minimum(deg[eu], deg[ev]) <= 8
    assert mid["n_light"] == int(light.sum()) and mid["n_light"] + mid["n_heavy"] == len(eu)
    np.testing.assert_array_equal(mid["supports"][light], all_exact["supports"][light])
    np.testing.assert_array_equal(mid["supports"][~light], all_hll["supports"][~light])
