"""Exact-count parity on the GPU against the CPU oracle (bit-exact integers)."""

import json
import math
import os

import numpy as np
import pytest

from oracle import cbind
from oracle import tc_oracle as O
from triangle_counting_hll_amd import gpu_graph, gpu_hybrid
from triangle_counting_hll_amd.graphgen import config_pairs, erdos_renyi, random_geometric, rmat

# oracle counts of the full-size graphs (tests/golden/make_full_size_counts.py)
with open(os.path.join(os.path.dirname(__file__), "golden", "full_size_counts.json")) as _f:
    FULL = json.load(_f)

pytestmark = pytest.mark.gpu


def _count(csr, **kw):
    return gpu_hybrid.gpu_exact_triangle_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v,
                                               **kw)[0]


def _truth(pairs):
    rp, col, *_ = O.csr_from_pairs(pairs)
    return cbind.exact_count(rp, col)


def test_reference_known_answer(device):
    # cpu_exp/triangle_hll_experiment.py:1054-1077: K4 + pendant path = 4
    csr = gpu_graph.build_csr_gpu(O.SELF_TEST_EDGES)
    assert _count(csr) == 4


@pytest.mark.parametrize("n", [3, 4, 10, 60])
def test_complete_graph_closed_form(device, n):
    iu = np.triu_indices(n, 1)
    pairs = np.stack(iu, axis=1).astype(np.int32)
    csr = gpu_graph.build_csr_gpu(pairs)
    assert _count(csr) == math.comb(n, 3)


@pytest.mark.parametrize("n", [5, 6, 7, 9, 17, 40])
def test_tail_units_rows_at_the_end_of_col(device, n):
    # K_n oriented by id: the last rows of col are 3, 2 and 1 ids long and,
    # with light_max 1, are streamed as the shorter row of heavy edges; a
    # masked last unit there would read past col's last id, so those rows
    # take the single-id path
    pairs = np.stack(np.triu_indices(n, 1), axis=1).astype(np.int32)
    csr = gpu_graph.build_csr_gpu(pairs)
    for params in ({"light_max": 1, "tail_units": True},
                   {"light_max": 1, "tail_units": True, "trim_rows": True},
                   {"light_max": 1, "tail_units": True, "plan": "atomic"},
                   {"light_max": 1, "tail_units": True, "trim_rows": True, "item_order": True,
                    "chunk": 1},
                   {"light_max": 1, "item_order": True, "block_min_bits": 8, "chunk": 3}):
        assert _count(csr, params=params) == math.comb(n, 3), params


@pytest.mark.parametrize("order", ["csr", "csr_mode", "shuffled", "reversed", "subset"])
def test_trim_search_any_edge_order(device, order):
    # with trim_rows the edge pass looks for the cut of u's row next to the
    # edge's own index (v's position in u's row in the CSR's edge order);
    # other orders give it a useless start and take the fallback searches,
    # so the count must not depend on the order.  subset: fewer edges than
    # twice the vertices, where the pass reads the rows' last ids alone and
    # searches whole rows; the count is the sum of those edges' supports
    import torch

    pairs = config_pairs("chunglu", 5)
    csr = gpu_graph.build_csr_gpu(pairs)
    eu, ev = csr.edge_u, csr.edge_v
    if order == "subset":
        keep = np.random.default_rng(9).permutation(eu.numel())[: csr.n_nodes]
        keep_d = torch.from_numpy(np.sort(keep)).to(eu.device)
        eu, ev = eu[keep_d], ev[keep_d]
        assert eu.numel() < 2 * csr.n_nodes
        truth = int(cbind.edge_supports(csr.row_ptr.cpu().numpy(), csr.col_idx.cpu().numpy(),
                                        eu.cpu().numpy(), ev.cpu().numpy()).sum())
    else:
        truth = _truth(pairs)
    if order == "shuffled":
        perm = torch.from_numpy(np.random.default_rng(7).permutation(eu.numel())).to(eu.device)
        eu, ev = eu[perm], ev[perm]
    elif order == "reversed":
        eu, ev = eu.flip(0), ev.flip(0)
    if order == "csr_mode":
        eu = ev = None
    for params in ({"trim_rows": True}, {"trim_rows": True, "light_max": 4, "tail_units": True},
                   {"trim_rows": True, "wide_offsets": True}):
        cnt = int(gpu_hybrid.exact_count(csr.row_ptr, csr.col_idx, eu, ev,
                                         params=params).item())
        assert cnt == truth, (order, params)


def test_clique_beyond_lds_table(device):
    # K_1400: forward rows of 1024..1399 ids exceed the per-wave LDS table and
    # are taken by whole workgroups over the joined table; C(1400, 3) triangles.
    n = 1400
    pairs = np.stack(np.triu_indices(n, 1), axis=1).astype(np.int32)
    csr = gpu_graph.build_csr_gpu(pairs)
    assert _count(csr) == math.comb(n, 3)


@pytest.mark.parametrize("plan,pack_ids", [("sorted", False), ("sorted", True), ("atomic", False)])
@pytest.mark.parametrize("wide", [False, True])
@pytest.mark.parametrize("n", [2100, 4400])
@pytest.mark.parametrize("trim,tail", [(False, False), (True, True), (False, True)])
def test_big_centers_mixed_with_small(device, plan, pack_ids, wide, n, trim, tail):
    # a K_n block plus an R-MAT graph on disjoint ids: forward rows of
    # 1024..4095 ids take the joined workgroup table, rows of 4096 ids and up
    # (n = 4400) the 128 KiB workgroup kernel, the R-MAT rows one wavefront
    # each; counts add up.  wide: 64-bit id addressing, which graphs above
    # 2^30 oriented edges take
    clique = np.stack(np.triu_indices(n, 1), axis=1)
    other = rmat(12, 16, seed=8).astype(np.int64) + n
    pairs = np.concatenate([clique, other]).astype(np.int32)
    csr = gpu_graph.build_csr_gpu(pairs)
    # trim: streamed rows cut to the center row's largest id by the edge pass;
    # tail: rows' last len % 4 ids in a masked 16-byte load
    assert (_count(csr, params={"plan": plan, "pack_ids": pack_ids, "wide_offsets": wide,
                                "trim_rows": trim, "tail_units": tail})
            == math.comb(n, 3) + _truth(other))


def test_rmat14_config(device):
    pairs = config_pairs("rmat14")
    rp, col, fdeg, nid, eu, ev, _ = O.csr_from_pairs(pairs)
    truth = O.exact_count_csr(rp, col)
    assert truth == cbind.exact_count(rp, col)
    csr = gpu_graph.build_csr_gpu(pairs)
    assert _count(csr) == truth
    # the reference's own int32 CSR handed in from the host
    got, secs = gpu_hybrid.gpu_exact_triangle_count(rp.astype(np.int32), col, eu, ev)
    assert got == truth and secs > 0


@pytest.mark.parametrize("plan,pack_ids", [("sorted", False), ("sorted", True), ("atomic", False)])
@pytest.mark.parametrize("light_max,chunk,wpc,unroll,blk", [(0, 1, 4, 1, 8), (2, 7, 8, 2, 13),
                                                            (16, 256, 16, 4, 0),
                                                            (64, 32, 32, 8, 10),
                                                            (32, 256, 16, 16, 10),
                                                            (10**6, 256, 16, 4, 9)])
@pytest.mark.parametrize("wide", [False, True])
@pytest.mark.parametrize("trim,tail", [(False, False), (True, True), (False, True)])
def test_tuning_knobs_do_not_change_count(device, light_max, chunk, wpc, unroll, blk, plan,
                                          pack_ids, wide, trim, tail):
    # pack_ids: the sorted plan's packed words carry the other's id instead of
    # its row start and length (the layout large graphs fall back to); wide:
    # 64-bit id addressing in the intersection kernels (graphs above 2^30
    # oriented edges); trim: the edge pass cuts each streamed row to its ids
    # at most the center row's largest id; tail: each streamed row's last
    # len % 4 ids read in a masked 16-byte load with its last whole ids
    pairs = rmat(12, 16, seed=7)
    truth = _truth(pairs)
    csr = gpu_graph.build_csr_gpu(pairs)
    params = {"light_max": light_max, "chunk": chunk, "waves_per_cu": wpc, "unroll": unroll,
              "plan": plan, "block_min_bits": blk, "pack_ids": pack_ids, "wide_offsets": wide,
              "trim_rows": trim, "tail_units": tail}
    assert _count(csr, params=params) == truth


@pytest.mark.parametrize("graph", ["rmat12", "chunglu", "clique_mix"])
@pytest.mark.parametrize("chunk,unroll,blk", [(1, 1, 8), (7, 2, 13), (64, 16, 10), (256, 4, 9),
                                              (1024, 16, 11)])
@pytest.mark.parametrize("layout", ["rows", "ids", "wide"])
def test_item_order_does_not_change_count(device, graph, chunk, unroll, blk, layout):
    # item_order: the wavefront phase's items sorted by the position of their
    # middle streamed row and handed out from one queue per XCD (the workgroup
    # phase and the 128 KiB kernel keep their items); every item is counted
    # once whatever the chunk, phase bounds and packed-word layout
    if graph == "rmat12":
        pairs = rmat(12, 16, seed=7)
    elif graph == "chunglu":
        pairs = config_pairs("chunglu", 6)
    else:
        n = 4400
        clique = np.stack(np.triu_indices(n, 1), axis=1)
        pairs = np.concatenate([clique, rmat(12, 16, seed=8).astype(np.int64) + n]).astype(np.int32)
    truth = _truth(pairs)
    csr = gpu_graph.build_csr_gpu(pairs)
    params = {"chunk": chunk, "unroll": unroll, "block_min_bits": blk, "item_order": True,
              "pack_ids": layout == "ids", "wide_offsets": layout == "wide",
              "trim_rows": layout == "rows", "tail_units": unroll >= 4}
    assert _count(csr, params=params) == truth, params


@pytest.mark.parametrize("blk", [8, 10, 13])
def test_block_phase_bounds_on_cliques(device, blk):
    # K_1400 rows of 128..1399 ids under every workgroup-phase lower bound,
    # including none (13: rows above 1023 then take the global search)
    n = 1400
    pairs = np.stack(np.triu_indices(n, 1), axis=1).astype(np.int32)
    csr = gpu_graph.build_csr_gpu(pairs)
    assert _count(csr, params={"block_min_bits": blk}) == math.comb(n, 3)
    with pytest.raises(ValueError, match="block_min_bits"):
        _count(csr, params={"block_min_bits": 7})
    with pytest.raises(ValueError, match="unroll"):
        _count(csr, params={"unroll": 3})


def test_center_beyond_workgroup_table_takes_global_search(device):
    # K_16385, the smallest graph with a forward row above 16383 ids (the
    # most the 128 KiB workgroup table holds): vertex 0 has d+ = 16384, so
    # its work items take the global binary search inside k_heavy_big
    import torch

    n = 16385
    iu = torch.triu_indices(n, n, 1, device="cuda")
    pairs = iu.t().to(torch.int32).contiguous()
    del iu
    csr = gpu_graph.build_csr_gpu(pairs)
    del pairs
    assert int(csr.degrees.max().item()) == n - 1
    assert _count(csr) == math.comb(n, 3)


def test_rgg_small(device):
    pairs = random_geometric(1 << 15, 2.8, seed=1)
    csr = gpu_graph.build_csr_gpu(pairs)
    assert _count(csr) == _truth(pairs)


def test_erdos_renyi(device):
    pairs = erdos_renyi(5000, 8 / 5000, seed=3)
    csr = gpu_graph.build_csr_gpu(pairs)
    assert _count(csr) == _truth(pairs)


def test_chunglu_downscaled(device):
    pairs = config_pairs("chunglu", scale_down=6)
    csr = gpu_graph.build_csr_gpu(pairs, row_ptr_dtype="int64")
    assert csr.row_ptr.dtype.itemsize == 8
    assert _count(csr) == _truth(pairs)


def test_edge_subset_and_order(device):
    import torch

    pairs = rmat(11, 8, seed=2)
    rp, col, fdeg, nid, eu, ev, _ = O.csr_from_pairs(pairs)
    rng = np.random.default_rng(0)
    sel = rng.permutation(len(eu))[: len(eu) // 3]
    truth = cbind.exact_count(rp, col, eu[sel], ev[sel])
    got, _ = gpu_hybrid.gpu_exact_triangle_count(rp, col, eu[sel], ev[sel])
    assert got == truth
    # reversed edge direction gives the same per-edge intersection
    got_rev, _ = gpu_hybrid.gpu_exact_triangle_count(rp, col, ev[sel], eu[sel])
    assert got_rev == truth
    # empty edge list
    z = torch.zeros(0, dtype=torch.int32, device="cuda")
    assert gpu_hybrid.gpu_exact_triangle_count(rp, col, z, z)[0] == 0


def test_coo_null_edges_and_launcher(device):
    import torch

    pairs = rmat(10, 8, seed=9)
    rp, col, fdeg, nid, eu, ev, _ = O.csr_from_pairs(pairs)
    truth = cbind.exact_count(rp, col)
    cnt = gpu_hybrid.exact_count(rp, col)  # edge list NULL -> CSR's own COO
    assert int(cnt.item()) == truth
    counts = torch.full((1,), 5, dtype=torch.int64, device="cuda")
    gpu_hybrid.exact_count_kernel[123, 256](rp.astype(np.int32), col, eu, ev, len(eu), counts)
    torch.cuda.synchronize()
    assert int(counts.item()) == truth + 5  # launch objects add into counts


def test_edge_supports_match_oracle(device):
    pairs = rmat(11, 16, seed=4)
    rp, col, fdeg, nid, eu, ev, _ = O.csr_from_pairs(pairs)
    sup = gpu_hybrid.edge_supports(rp, col, eu, ev).cpu().numpy()
    assert np.array_equal(sup, cbind.edge_supports(rp, col, eu, ev))
    assert np.array_equal(sup, O.edge_supports(rp, col, eu, ev))


def test_repeat_is_deterministic(device):
    pairs = rmat(13, 8, seed=1)
    csr = gpu_graph.build_csr_gpu(pairs)
    a = [_count(csr) for _ in range(3)]
    assert len(set(a)) == 1 and a[0] == _truth(pairs)


@pytest.mark.parametrize("n", [0, 1, 15, 16, 17, 1023, 1024, 1025, 70_001])
@pytest.mark.parametrize("shift", [0, 1, 13])
def test_flag_indices_match_flatnonzero(device, n, shift):
    # chunk (1024 flags per wavefront) and 16-byte lane boundaries, 16-byte
    # aligned arrays (vector loads) and unaligned slices (byte loads), flag
    # values other than 1 (any nonzero byte counts)
    import torch

    rng = np.random.default_rng(n + shift)
    host = (rng.random(n + shift) < 0.3).astype(np.uint8) * rng.integers(1, 256, n + shift,
                                                                          dtype=np.uint8)
    dev = torch.from_numpy(host).cuda()[shift:]
    got = gpu_hybrid.flag_indices(dev).cpu().numpy()
    assert np.array_equal(got, np.flatnonzero(host[shift:]))


def test_flag_indices_all_set_and_none_set(device):
    import torch

    for fill in (0, 1):
        f = torch.full((5000,), fill, dtype=torch.uint8, device="cuda")
        got = gpu_hybrid.flag_indices(f).cpu().numpy()
        assert np.array_equal(got, np.flatnonzero(np.full(5000, fill)))


def test_chunglu_full_size_matches_cpu_oracle(device):
    # the configured 117M-edge power-law graph (n = 3,072,441, beta 2.1,
    # w_max 33,000, seed 2) with int64 offsets: the count of both packed-word
    # layouts and of the atomic plan equals the C oracle on the same CSR
    csr = gpu_graph.build_csr_gpu(config_pairs("chunglu"), row_ptr_dtype="int64")
    rp = csr.row_ptr.cpu().numpy()
    col = csr.col_idx.cpu().numpy()
    truth = cbind.exact_count(rp, col)
    assert truth == FULL["chunglu"]["triangles"]
    for params in ({}, {"pack_ids": True}, {"plan": "atomic"}, {"wide_offsets": True},
                   {"trim_rows": True}, {"trim_rows": True, "wide_offsets": True},
                   {"trim_rows": True, "tail_units": True},
                   {"tail_units": True, "wide_offsets": True, "plan": "atomic"},
                   {"trim_rows": True, "tail_units": True, "item_order": True, "chunk": 256},
                   {"item_order": True, "block_min_bits": 9, "pack_ids": True},
                   # the knobs the bench's tuner picks on this graph
                   {"light_max": 44, "chunk": 128, "block_min_bits": 11, "trim_rows": True,
                    "tail_units": True, "item_order": True}):
        assert _count(csr, params=params) == truth, params


def test_rmat23_full_size_planners_agree(device):
    # the 129M-edge R-MAT scale 23 graph (seed 4) built and counted on the
    # device: the planners and packed-word layouts agree with each other and
    # with the C oracle's count of this graph (golden file)
    from triangle_counting_hll_amd import batch

    csr = gpu_graph.build_csr_gpu(batch.rmat_gpu(23, 16, seed=4), row_ptr_dtype="int64")
    assert csr.n_oriented == FULL["rmat23"]["oriented_edges"]
    counts = {str(p): _count(csr, params=p)
              for p in ({}, {"pack_ids": True}, {"plan": "atomic"}, {"unroll": 4, "chunk": 64},
                        {"wide_offsets": True}, {"trim_rows": True}, {"tail_units": True},
                        {"item_order": True, "chunk": 256}, {"item_order": True, "unroll": 2})}
    assert set(counts.values()) == {FULL["rmat23"]["triangles"]}, counts


def test_rgg21_full_size_matches_cpu_oracle(device):
    # the configured random geometric graph (n = 2^21, mean degree 2.8, seed
    # 1): every edge is merged in the classify pass; eager (early exit after
    # the merge pass) and the id layout agree with the C oracle
    csr = gpu_graph.build_csr_gpu(config_pairs("rgg21"))
    rp = csr.row_ptr.cpu().numpy()
    col = csr.col_idx.cpu().numpy()
    truth = cbind.exact_count(rp, col)
    assert truth == _truth(config_pairs("rgg21"))
    for params in ({}, {"pack_ids": True}, {"light_max": 2},
                   {"light_max": 2, "item_order": True, "trim_rows": True, "chunk": 128}):
        assert _count(csr, params=params) == truth, params
