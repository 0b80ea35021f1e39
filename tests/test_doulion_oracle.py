"""DOULION oracle (experiment/exp_gpu.py:396-450) pinned by numpy's own generator,
the q = 1 closed form, re-orientation invariance and unbiasedness."""

import numpy as np
import pytest

from oracle import tc_oracle as O
from triangle_counting_hll_amd import graphgen


@pytest.mark.parametrize("seed", [0, 42, 51, 2**40 + 3])
def test_pcg64_restatement_matches_numpy(seed):
    rng = np.random.default_rng(seed)
    st = rng.bit_generator.state["state"]
    assert np.array_equal(O.pcg64_doubles(st["state"], st["inc"], 700), rng.random(700))


def test_q_one_keeps_every_edge():
    rp, col, *_ = O.csr_from_pairs(graphgen.rmat(10, 8, seed=1))
    est, cnt = O.doulion_forward(rp, col, 1.0, n_seeds=2)
    assert cnt == [O.exact_count_csr(rp, col)] * 2 and est == [float(c) for c in cnt]


def test_sample_count_is_orientation_free():
    # counting the sample with the original orientation equals counting the
    # sampled undirected edges re-oriented from scratch (src/doulion.py:41)
    rp, col, fdeg, nid, eu, ev, _ = O.csr_from_pairs(graphgen.rmat(10, 16, seed=2))
    q, seed = 0.3, 44
    _, cnt = O.doulion_forward(rp, col, q, n_seeds=1, base_seed=seed)
    mask = np.random.default_rng(seed).random(len(col)) < q
    sampled = np.stack([eu[mask], ev[mask]], 1)
    rp2, col2, *_ = O.csr_from_pairs(sampled)
    assert cnt[0] == O.exact_count_csr(rp2, col2)


def test_estimator_is_unbiased():
    n = 40
    rp, col, *_ = O.csr_from_pairs(np.stack(np.triu_indices(n, 1), axis=1))
    T = n * (n - 1) * (n - 2) // 6
    est, _ = O.doulion_forward(rp, col, 0.5, n_seeds=200, base_seed=1)
    # std of one estimate is below T (each triangle survives w.p. 1/8, scaled by 8)
    assert abs(np.mean(est) - T) < 5 * T / np.sqrt(200)
