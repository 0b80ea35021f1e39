"""Device CSR build / orientation / degree buckets against the CPU oracle (bit-exact)."""

import numpy as np
import pytest

from oracle import tc_oracle as O
from triangle_counting_hll_amd import gpu_graph
from triangle_counting_hll_amd.graphgen import config_pairs, random_geometric, rmat

pytestmark = pytest.mark.gpu


def _check(pairs, dtype="auto"):
    rp, col, fdeg, nid, eu, ev, udeg = O.csr_from_pairs(pairs)
    csr = gpu_graph.build_csr_gpu(pairs, row_ptr_dtype=dtype)
    assert csr.n_nodes == len(rp) - 1 and csr.n_oriented == len(col)
    assert np.array_equal(csr.row_ptr.cpu().numpy().astype(np.int64), rp)
    assert np.array_equal(csr.col_idx.cpu().numpy(), col)
    assert np.array_equal(csr.degrees.cpu().numpy(), fdeg)
    assert np.array_equal(csr.node_ids.cpu().numpy(), nid)
    assert np.array_equal(csr.edge_u.cpu().numpy(), eu)
    assert np.array_equal(csr.edge_v.cpu().numpy(), ev)
    assert np.array_equal(csr.undirected_deg.cpu().numpy(), udeg)
    return csr


def test_self_test_graph(device):
    _check(O.SELF_TEST_EDGES)


def test_rmat14_with_duplicates_both_orientations_and_loops(device):
    pairs = config_pairs("rmat14")
    rng = np.random.default_rng(5)
    extra = pairs[rng.integers(0, len(pairs), 5000)][:, ::-1]
    loops = np.repeat(rng.integers(0, 1 << 14, 300)[:, None], 2, axis=1).astype(np.int32)
    _check(np.concatenate([pairs, extra, loops]).astype(np.int32))


def test_sparse_ids_are_relabelled(device):
    pairs = rmat(10, 8, seed=3).astype(np.int64) * 7 + 3
    _check(pairs.astype(np.int32))


def test_rgg(device):
    _check(random_geometric(1 << 16, 2.8, seed=1), dtype="int32")


def test_long_rows_take_lds_and_global_sort(device):
    # a clique of 9000 on top of a sparse graph: forward rows up to 8999 ids
    # (beyond the LDS sort tier) plus rows of every size below
    n = 9000
    iu = np.triu_indices(n, 1)
    rng = np.random.default_rng(0)
    keep = rng.random(len(iu[0])) < 0.25
    clique = np.stack([iu[0][keep], iu[1][keep]], axis=1)
    rest = rmat(12, 8, seed=2) + n
    star = np.stack([np.zeros(20000, dtype=np.int64), np.arange(20000) + 20000], axis=1)
    pairs = np.concatenate([clique, rest, star]).astype(np.int32)
    csr = _check(pairs, dtype="int64")
    assert int((csr.row_ptr[1:] - csr.row_ptr[:-1]).max()) > 64


def test_dense_clique_global_tier(device):
    n = 9000
    pairs = np.stack(np.triu_indices(n, 1), axis=1)[::-1].astype(np.int32)
    csr = _check(pairs)
    assert int(csr.degrees.max()) == n - 1


def test_empty_and_all_loops(device):
    csr = gpu_graph.build_csr_gpu(np.zeros((0, 2), dtype=np.int32))
    assert csr.n_nodes == 0 and csr.n_oriented == 0 and csr.row_ptr.numel() == 1
    csr = gpu_graph.build_csr_gpu(np.array([[3, 3], [5, 5]], dtype=np.int32))
    assert csr.n_nodes == 0 and csr.n_oriented == 0


def test_negative_ids_raise(device):
    with pytest.raises(ValueError):
        gpu_graph.build_csr_gpu(np.array([[0, 1], [-2, 3]], dtype=np.int32))


def test_to_csr_gpu_matches_reference_layout(device):
    pairs = rmat(9, 8, seed=4)
    edges = {(int(min(a, b)), int(max(a, b))) for a, b in pairs if a != b}
    fwd, _ = O.orient_graph(edges)
    ref = O.to_csr_numpy(fwd)
    got = gpu_graph.to_csr_gpu(fwd)
    for a, b in zip(ref, got):
        assert np.array_equal(a, b.cpu().numpy())


def test_degree_buckets(device):
    pairs = config_pairs("rmat14")
    rp, *_ = O.csr_from_pairs(pairs)
    starts, perm = gpu_graph.degree_buckets(rp.astype(np.int32))
    deg = np.diff(rp)
    bucket = np.where(deg == 0, 0, np.floor(np.log2(np.maximum(deg, 1))).astype(int) + 1)
    order = np.lexsort((np.arange(len(deg)), bucket))
    assert np.array_equal(perm.cpu().numpy(), order)
    counts = np.bincount(bucket, minlength=33)
    assert np.array_equal(starts.cpu().numpy(), np.concatenate([[0], np.cumsum(counts)]))
