"""DOULION on the GPU: sampling mask bit-exact against numpy's generator, sampled
CSR and estimates equal to the oracle's."""

import numpy as np
import pytest

from oracle import cbind
from oracle import tc_oracle as O
from triangle_counting_hll_amd import doulion, gpu_graph, graphgen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 1000, 12_345_677])
@pytest.mark.parametrize("q", [0.05, 0.5, 1.0])
def test_mask_matches_numpy_generator(device, n, q):
    rng = np.random.default_rng(42)
    got = doulion.sample_mask_gpu(n, q, rng).cpu().numpy()
    want = np.random.default_rng(42).random(n) < q
    assert np.array_equal(got.astype(bool), want)
    # the host generator moved past the n draws the device consumed
    ref = np.random.default_rng(42)
    ref.random(n)
    assert rng.random() == ref.random()


def test_mask_continues_a_used_generator(device):
    rng = np.random.default_rng(7)
    rng.random(1001)
    got = doulion.sample_mask_gpu(5000, 0.25, rng).cpu().numpy()
    ref = np.random.default_rng(7)
    ref.random(1001)
    assert np.array_equal(got.astype(bool), ref.random(5000) < 0.25)


def test_edge_subgraph_matches_numpy(device):
    import torch

    rp, col, *_ = O.csr_from_pairs(graphgen.rmat(12, 8, seed=3))
    keep = (np.random.default_rng(0).random(len(col)) < 0.3)
    srp, scol = doulion.edge_subgraph(rp.astype(np.int32), col,
                                      torch.from_numpy(keep.astype(np.uint8)).cuda())
    eu = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    want_rp = np.zeros(len(rp), np.int64)
    np.cumsum(np.bincount(eu[keep], minlength=len(rp) - 1), out=want_rp[1:])
    assert np.array_equal(srp.cpu().numpy(), want_rp)
    assert np.array_equal(scol.cpu().numpy(), col[keep])


@pytest.mark.parametrize("q", [0.1, 0.05, 0.5])
def test_estimates_equal_oracle(device, q):
    rp, col, *_ = O.csr_from_pairs(graphgen.rmat(12, 16, seed=5))
    res = doulion.doulion_forward_gpu(rp.astype(np.int32), col, q, n_seeds=4, base_seed=42)
    want, _ = O.doulion_forward(rp, col, q, n_seeds=4, base_seed=42)
    assert res["estimates"] == want
    assert res["mean_estimate"] == float(np.mean(want)) and len(res["times"]) == 4


def test_chunglu_sample_counts(device):
    # full-size Chung-Lu graph, q = 0.1 (the reference's first DOULION rate),
    # one seed: the sampled count equals the C oracle's on numpy's mask
    csr = gpu_graph.build_csr_gpu(graphgen.config_pairs("chunglu"), row_ptr_dtype="int64")
    est, secs, kept = doulion.doulion_gpu(csr.row_ptr, csr.col_idx, 0.1,
                                          np.random.default_rng(42))
    rp = csr.row_ptr.cpu().numpy()
    col = csr.col_idx.cpu().numpy()
    mask = np.random.default_rng(42).random(len(col)) < 0.1
    eu = np.repeat(np.arange(len(rp) - 1, dtype=np.int32), np.diff(rp))
    srp = np.zeros(len(rp), np.int64)
    np.cumsum(np.bincount(eu[mask], minlength=len(rp) - 1), out=srp[1:])
    want = cbind.exact_count(srp, col[mask])
    assert kept == int(mask.sum()) and est == want / 0.1 ** 3 and secs > 0
