"""Batched many-graph counting on the GPU: per-graph counts bit-exact against the oracle."""

import ctypes

import numpy as np
import pytest

from oracle import cbind
from oracle import tc_oracle as O
from triangle_counting_hll_amd import _lib, batch, gpu_hybrid
from triangle_counting_hll_amd.graphgen import rmat

pytestmark = pytest.mark.gpu


def _per_graph_truth(spec, pairs):
    """Oracle count of every graph of the batch, each built on its own."""
    out = []
    po = spec.pair_off
    for g in range(spec.n_graphs):
        p = pairs[po[g]:po[g + 1]]
        if len(p) == 0:
            out.append(0)
            continue
        rp, col, *_ = O.csr_from_pairs(p)
        out.append(cbind.exact_count(rp, col))
    return np.array(out, dtype=np.int64)


def test_device_generator_matches_host_twin(device):
    spec = batch.many_graphs_spec(48, seed=11, scale_down=5)
    host = batch.generate_pairs_numpy(spec)
    dev = batch.generate_pairs_gpu(spec).cpu().numpy()
    assert np.array_equal(host, dev)
    # one graph of the batch equals the standalone generator on its stream
    g = int(np.nonzero(spec.kind == 1)[0][0])
    lo, hi = spec.pair_off[g], spec.pair_off[g + 1]
    alone = rmat(int(spec.param[g]), 16, seed=11, stream=g)
    assert np.array_equal(dev[lo:hi] - spec.id_off[g], alone)


@pytest.mark.parametrize("plan", ["sorted", "atomic"])
def test_batch_counts_bit_exact(device, plan):
    spec = batch.many_graphs_spec(96, seed=3, scale_down=4)
    pairs = batch.generate_pairs_numpy(spec)
    truth = _per_graph_truth(spec, pairs)
    b = batch.build_batch(spec)
    got = batch.batch_exact_count(b, {"plan": plan}).cpu().numpy()
    assert np.array_equal(got, truth)
    plain, secs, dev_ms = batch.plain_per_graph_counts(b)
    assert np.array_equal(plain, truth) and secs > 0 and dev_ms > 0
    # the per-graph CSRs hold each graph rebased to ids from 0
    graphs = batch.per_graph_csrs(b)
    assert len(graphs) == b.n_graphs
    for g in (0, b.n_graphs // 2, b.n_graphs - 1):
        rp, col, *_ = O.csr_from_pairs(pairs[spec.pair_off[g]:spec.pair_off[g + 1]])
        assert np.array_equal(graphs[g].row_ptr.cpu().numpy(), rp)
        assert np.array_equal(graphs[g].col_idx.cpu().numpy(), col)


@pytest.mark.parametrize("plan", ["sorted", "atomic"])
def test_batch_with_workgroup_centers(device, plan):
    # cliques whose forward rows reach 1024..4095 ids (joined workgroup table)
    # and >= 4096 ids (128 KiB workgroup kernel) next to small R-MAT graphs:
    # each graph's count lands in its own slot
    import math

    import torch

    parts = [np.stack(np.triu_indices(1500, 1), axis=1), rmat(10, 16, seed=2),
             np.stack(np.triu_indices(4300, 1), axis=1), rmat(9, 8, seed=3)]
    n_ids = np.array([1500, 1 << 10, 4300, 1 << 9], dtype=np.int64)
    spec = batch.BatchSpec(kind=np.zeros(4, np.uint8), param=np.zeros(4, np.int32), n_ids=n_ids,
                           n_pairs=np.array([len(p) for p in parts], dtype=np.int64),
                           keys=np.zeros(4, np.uint64), seed=0)
    pairs = np.concatenate([p.astype(np.int64) + off
                            for p, off in zip(parts, spec.id_off[:-1])]).astype(np.int32)
    truth = np.array([math.comb(1500, 3), cbind.exact_count(*O.csr_from_pairs(parts[1])[:2]),
                      math.comb(4300, 3), cbind.exact_count(*O.csr_from_pairs(parts[3])[:2])],
                     dtype=np.int64)
    b = batch.build_batch(spec, pairs=torch.from_numpy(pairs).cuda())
    got = batch.batch_exact_count(b, {"plan": plan}).cpu().numpy()
    assert np.array_equal(got, truth)


def _clique_and_rmat_batch(parts, n_ids):
    import torch

    n_ids = np.asarray(n_ids, dtype=np.int64)
    G = len(parts)
    spec = batch.BatchSpec(kind=np.zeros(G, np.uint8), param=np.zeros(G, np.int32), n_ids=n_ids,
                           n_pairs=np.array([len(p) for p in parts], dtype=np.int64),
                           keys=np.zeros(G, np.uint64), seed=0)
    pairs = np.concatenate([p.astype(np.int64) + off
                            for p, off in zip(parts, spec.id_off[:-1])]).astype(np.int32)
    truth = np.array([cbind.exact_count(*O.csr_from_pairs(p)[:2]) for p in parts], np.int64)
    return batch.build_batch(spec, pairs=torch.from_numpy(pairs).cuda()), truth


@pytest.mark.parametrize("shuffle", [False, True])
def test_batch_sort_width_follows_the_layout(device, shuffle):
    # Edges grouped by graph in CSR order let the planner sort the heavy
    # edges by the low bits of their center only (here 7 bits, one radix
    # pass: the largest graph has 65..128 vertices, the batch about 11,900,
    # 14 bits); with the edge order shuffled the layout check fails and every
    # center bit is sorted (two passes).  Graph sizes sit on both sides of
    # powers of two (64, 65, 128), where centers of neighbouring graphs share
    # their low bits.
    import torch

    parts = [np.stack(np.triu_indices(64, 1), axis=1), np.stack(np.triu_indices(65, 1), axis=1),
             rmat(7, 16, seed=4), np.stack(np.triu_indices(40, 1), axis=1)]
    parts = parts * 40
    b, truth = _clique_and_rmat_batch(parts, [64, 65, 128, 40] * 40)
    eu, ev = b.csr.edge_u, b.csr.edge_v
    if shuffle:
        perm = torch.from_numpy(np.random.default_rng(1).permutation(b.csr.n_oriented)).cuda()
        eu, ev = eu[perm].contiguous(), ev[perm].contiguous()
    counts = torch.zeros(b.n_graphs, dtype=torch.int64, device="cuda")
    rp, is64 = _lib.row_ptr_arg(b.csr.row_ptr)
    for params in ({}, {"pack_ids": True}, {"trim_rows": True}, {"tail_units": True}):
        counts.zero_()
        p = gpu_hybrid._exact_params(params)
        _lib.check(_lib.load().tc_batch_exact_count(
            _lib.ptr(rp), is64, _lib.ptr(b.csr.col_idx), b.csr.n_nodes, _lib.ptr(eu),
            _lib.ptr(ev), b.csr.n_oriented, _lib.ptr(b.graph_of), b.n_graphs, _lib.ptr(counts),
            ctypes.byref(p) if p is not None else None, _lib.stream_ptr()))
        assert np.array_equal(counts.cpu().numpy(), truth), params


def test_graph_replays_match_eager(device):
    import torch

    spec = batch.many_graphs_spec(64, seed=5, scale_down=4)
    b = batch.build_batch(spec)
    eager = batch.batch_exact_count(b).cpu().numpy()
    g = batch.BatchCountGraph(b, {"chunk": 128, "unroll": 4})
    assert g.workspace_bytes > 0
    for i in range(4):
        # unrelated eager work and synchronisation between replays
        x = torch.arange(1000, device="cuda").sum()
        torch.cuda.synchronize()
        got = g.launch().clone()
        torch.cuda.synchronize()
        assert np.array_equal(got.cpu().numpy(), eager), f"replay {i}"
        assert int(x.item()) == 499500
    g.close()


def test_empty_graphs_in_batch(device):
    spec = batch.many_graphs_spec(16, seed=2, scale_down=4)
    spec.n_pairs[3] = 0
    spec.n_pairs[7] = 0
    pairs = batch.generate_pairs_numpy(spec)
    truth = _per_graph_truth(spec, pairs)
    b = batch.build_batch(spec)
    got = batch.batch_exact_count(b).cpu().numpy()
    assert np.array_equal(got, truth)
    assert got[3] == 0 and got[7] == 0


def test_graph_replays_with_only_light_edges(device):
    # every edge is merged in the classify pass (rows of at most 3 ids), and a
    # captured pipeline never takes the eager early exit: the radix sort drops
    # every word and the intersection kernels get no work item
    import math

    import torch

    parts = [np.stack(np.triu_indices(4, 1), axis=1), np.array([[0, 1], [1, 2], [2, 0]]),
             np.array([[0, 1], [1, 2], [2, 3], [3, 4]]), np.stack(np.triu_indices(3, 1), axis=1)]
    n_ids = np.array([4, 3, 5, 3], dtype=np.int64)
    spec = batch.BatchSpec(kind=np.zeros(4, np.uint8), param=np.zeros(4, np.int32), n_ids=n_ids,
                           n_pairs=np.array([len(p) for p in parts], dtype=np.int64),
                           keys=np.zeros(4, np.uint64), seed=0)
    pairs = np.concatenate([p.astype(np.int64) + off
                            for p, off in zip(parts, spec.id_off[:-1])]).astype(np.int32)
    truth = np.array([math.comb(4, 3), 1, 0, 1], dtype=np.int64)
    b = batch.build_batch(spec, pairs=torch.from_numpy(pairs).cuda())
    assert np.array_equal(batch.batch_exact_count(b).cpu().numpy(), truth)
    g = batch.BatchCountGraph(b)
    for i in range(3):
        got = g.launch().clone()
        torch.cuda.synchronize()
        assert np.array_equal(got.cpu().numpy(), truth), f"replay {i}"
    g.close()


def test_batch4096_full_size(device):
    # the configured many-graph workload (4,096 graphs, seed 3: half ER with n
    # in [200, 50,000] and p = 8/n, half R-MAT scale 8..16, edge factor 16),
    # 607M oriented edges in one packed CSR: the eager pipeline and two
    # replays of the captured launch give every graph the count the C oracle
    # finds on that graph's edge range of the same CSR
    import torch

    from triangle_counting_hll_amd import autotune

    spec = batch.many_graphs_spec()
    assert spec.n_graphs == 4096
    b = batch.build_batch(spec)
    rp = b.csr.row_ptr.cpu().numpy().astype(np.int64)
    col = b.csr.col_idx.cpu().numpy()
    eu = b.csr.edge_u.cpu().numpy()
    truth = np.array([cbind.exact_count(rp, col, eu[b.edge_off[g]:b.edge_off[g + 1]],
                                        col[b.edge_off[g]:b.edge_off[g + 1]])
                      for g in range(b.n_graphs)], dtype=np.int64)
    assert (truth > 0).sum() > 4000
    eager = batch.batch_exact_count(b).cpu().numpy()
    assert np.array_equal(eager, truth)
    params = dict(autotune.DEFAULTS, chunk=256, block_min_bits=9)
    g = batch.BatchCountGraph(b, params)
    for i in range(2):
        x = torch.arange(1000, device="cuda").sum()
        torch.cuda.synchronize()
        got = g.launch().clone()
        torch.cuda.synchronize()
        assert np.array_equal(got.cpu().numpy(), truth), f"replay {i}"
        assert int(x.item()) == 499500
    g.close()


def test_batch_plan_falls_back_on_other_layouts(device):
    # The planner narrows its sort and packed words and orders work by graph
    # only when the batch's layout allows it (graph ids ascending along the
    # vertices, no edge joining two graphs, edges grouped by graph); any other
    # layout takes the single-graph form, and the counts stay exact.
    import torch

    spec = batch.many_graphs_spec(64, seed=7, scale_down=4)
    pairs = batch.generate_pairs_numpy(spec)
    truth = _per_graph_truth(spec, pairs)
    b = batch.build_batch(spec)
    rp, is64 = _lib.row_ptr_arg(b.csr.row_ptr)

    def run(bb, graph_of, G):
        counts = torch.zeros(G, dtype=torch.int64, device="cuda")
        _lib.check(_lib.load().tc_batch_exact_count(
            _lib.ptr(_lib.row_ptr_arg(bb.csr.row_ptr)[0]), is64, _lib.ptr(bb.csr.col_idx),
            bb.csr.n_nodes, _lib.ptr(bb.csr.edge_u), _lib.ptr(bb.csr.edge_v), bb.csr.n_oriented,
            _lib.ptr(graph_of), G, _lib.ptr(counts), None, _lib.stream_ptr()))
        return counts.cpu().numpy()

    assert np.array_equal(run(b, b.graph_of, b.n_graphs), truth)
    # graph ids descending along the vertices
    rev = (b.n_graphs - 1 - b.graph_of).to(torch.int32).contiguous()
    assert np.array_equal(run(b, rev, b.n_graphs)[::-1], truth)
    # the whole batch as one graph
    one = torch.zeros_like(b.graph_of)
    assert run(b, one, 1)[0] == truth.sum()
    # edges that join graphs: the total is the whole graph's count, and the
    # graphs no such edge touches keep their own counts
    rng = np.random.default_rng(3)
    g1 = rng.choice(b.n_graphs, 6, replace=False)
    g2 = (g1 + 1) % b.n_graphs
    off = spec.id_off
    cross = np.stack([off[g1] + rng.integers(0, spec.n_ids[g1]),
                      off[g2] + rng.integers(0, spec.n_ids[g2])], axis=1)
    both = np.concatenate([pairs, cross.astype(np.int32)])
    bx = batch.build_batch(spec, pairs=torch.from_numpy(both).cuda())
    got = run(bx, bx.graph_of, bx.n_graphs)
    rpx = bx.csr.row_ptr.cpu().numpy().astype(np.int64)
    colx = bx.csr.col_idx.cpu().numpy()
    assert got.sum() == cbind.exact_count(rpx, colx)
    untouched = np.setdiff1d(np.arange(b.n_graphs), np.concatenate([g1, g2]))
    assert np.array_equal(got[untouched], truth[untouched])
