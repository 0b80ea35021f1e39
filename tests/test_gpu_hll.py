"""HLL registers (bit-exact), FP64 estimates (<= 1e-12) and hybrid parity on the GPU."""

import math

import numpy as np
import pytest

from oracle import tc_oracle as O
from triangle_counting_hll_amd import gpu_graph, gpu_hll, gpu_hybrid
from triangle_counting_hll_amd.graphgen import config_pairs, rmat

pytestmark = pytest.mark.gpu

GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def test_device_hash_known_answers(device):
    # published splitmix64 outputs for seed 0 (state k * golden, k = 1..3); mix(0) = 0
    x = np.array([0, 0x9E3779B97F4A7C15, 0x3C6EF372FE94F82A, 0xDAA66D2C7DDF743F],
                 dtype=np.uint64).view(np.int64)
    h = gpu_hll.splitmix64(x).cpu().numpy().view(np.uint64)
    assert [int(v) for v in h] == [0, 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                   0x06C45D188009454F]
    r = np.random.default_rng(0).integers(0, 2**63, 4096, dtype=np.int64)
    assert np.array_equal(gpu_hll.splitmix64(r).cpu().numpy().view(np.uint64), O.hash64(r))


@pytest.mark.parametrize("p,theta", [(10, 0), (4, 0), (8, 5), (12, 3), (16, 20)])
def test_registers_bit_exact(device, p, theta):
    pairs = config_pairs("rmat14")
    rp, col, *_ = O.csr_from_pairs(pairs)
    csr = gpu_graph.build_csr_gpu(pairs)
    n = csr.n_nodes
    regs, has = gpu_hll.alloc_sketch_arrays_device(n, p)
    gpu_hll.build_sketches_kernel[(n + 255) // 256, 256](csr.row_ptr, csr.col_idx, n, theta, p,
                                                          regs, has)
    ro, ho = O.build_registers(rp, col, p, theta)
    assert np.array_equal(has.cpu().numpy(), ho)
    assert np.array_equal(regs.cpu().numpy().reshape(n, -1), ro)


def test_numpy_arguments_are_written_back(device):
    pairs = rmat(10, 8, seed=1)
    rp, col, *_ = O.csr_from_pairs(pairs)
    n = len(rp) - 1
    regs, has = gpu_hll.alloc_sketch_arrays(n, 8)
    gpu_hll.build_sketches_kernel[1, 256](rp.astype(np.int32), col, n, 4, 8, regs, has)
    ro, ho = O.build_registers(rp, col, 8, 4)
    assert np.array_equal(regs.reshape(n, -1), ro) and np.array_equal(has, ho)


def test_accumulate_and_slot_layout(device):
    import torch

    pairs = rmat(10, 8, seed=6)
    rp, col, *_ = O.csr_from_pairs(pairs)
    n = len(rp) - 1
    p = 9
    m = 1 << p
    # accumulate: a preset row keeps its larger bytes (in-place max semantics)
    pre = np.zeros((n, m), dtype=np.uint8)
    pre[3, :] = 7
    regs = torch.from_numpy(pre.reshape(-1).copy()).cuda()
    has = torch.zeros(n, dtype=torch.uint8, device="cuda")
    gpu_hll.build_sketches(rp, col, n, 0, p, regs, has, accumulate=True)
    ro, _ = O.build_registers(rp, col, p, 0)
    assert np.array_equal(regs.cpu().numpy().reshape(n, m), np.maximum(ro, pre))
    gpu_hll.build_sketches(rp, col, n, 0, p, regs, has, accumulate=False)
    assert np.array_equal(regs.cpu().numpy().reshape(n, m), ro)
    # compact slots: only sketched vertices own a row
    theta = 6
    deg = np.diff(rp)
    sk = np.nonzero(deg >= theta)[0]
    slot = np.full(n, -1, dtype=np.int32)
    slot[sk] = np.arange(len(sk), dtype=np.int32)
    regs_c = torch.zeros(len(sk) * m, dtype=torch.uint8, device="cuda")
    gpu_hll.build_sketches(rp, col, n, theta, p, regs_c, has, slot=slot)
    ro, ho = O.build_registers(rp, col, p, theta)
    assert np.array_equal(regs_c.cpu().numpy().reshape(-1, m), ro[sk])
    assert np.array_equal(has.cpu().numpy(), ho)


def test_bad_precision_raises(device):
    regs, has = gpu_hll.alloc_sketch_arrays_device(4, 4)
    rp = np.array([0, 1, 2, 2, 2], dtype=np.int32)
    col = np.array([1, 2], dtype=np.int32)
    with pytest.raises(ValueError):
        gpu_hll.build_sketches(rp, col, 4, 0, 17, regs, has)
    with pytest.raises(ValueError):
        gpu_hll.union_estimate_pairs(regs, [0], [1], 3)


@pytest.mark.parametrize("p", [4, 5, 6, 8, 10, 12, 14, 16])
def test_union_estimates_fp64(device, p):
    pairs = config_pairs("rmat14")
    rp, col, *_ = O.csr_from_pairs(pairs)
    n = len(rp) - 1
    ro, _ = O.build_registers(rp, col, p, 0)
    rng = np.random.default_rng(p)
    us = rng.integers(0, n, 3000).astype(np.int32)
    vs = rng.integers(0, n, 3000).astype(np.int32)
    vs[:500] = us[:500]  # single-sketch estimates
    got = gpu_hll.union_estimate_pairs(ro.reshape(-1), us, vs, p).cpu().numpy()
    want = O.union_estimate_rows(ro[us], ro[vs])
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)
    # the per-row restatement itself agrees with the scalar one
    for i in range(0, 3000, 301):
        assert math.isclose(want[i], O.estimate_from_registers(np.maximum(ro[us[i]], ro[vs[i]])),
                            rel_tol=1e-15)


def test_estimator_branches_on_synthetic_registers(device):
    # all-zero rows (E = 0), constant rows, LinearCounting rows and the
    # large-range correction (rows of 58 at p = 4 give E > 2^64 / 30)
    p = 4
    m = 1 << p
    rows = np.zeros((6, m), dtype=np.uint8)
    rows[1, :] = 3
    rows[2, : m // 2] = 1
    rows[3, :] = 58
    rows[4, :] = 61
    rows[5, ::3] = 2
    idx = np.arange(6, dtype=np.int32)
    got = gpu_hll.union_estimate_pairs(rows.reshape(-1), idx, idx, p).cpu().numpy()
    want = O.estimate_rows(rows)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)
    assert got[0] == 0.0
    assert math.isclose(got[1], 0.673 * m * 8, rel_tol=1e-15)


def _hybrid_setup(p=10, theta=0, scale=None):
    pairs = config_pairs("rmat14") if scale is None else rmat(scale, 16, seed=11)
    rp, col, fdeg, nid, eu, ev, _ = O.csr_from_pairs(pairs)
    csr = gpu_graph.build_csr_gpu(pairs)
    n = csr.n_nodes
    regs, has = gpu_hll.alloc_sketch_arrays_device(n, p)
    gpu_hll.build_sketches_kernel[1, 256](csr.row_ptr, csr.col_idx, n, theta, p, regs, has)
    ro, ho = O.build_registers(rp, col, p, theta)
    return rp, col, eu, ev, csr, regs, has, ro, ho


@pytest.mark.parametrize("gamma,theta", [(0, 0), (8, 0), (8, 12), (40, 0), (10**9, 0)])
@pytest.mark.parametrize("split", [True, False])
def test_hybrid_matches_oracle(device, gamma, theta, split):
    p = 10
    rp, col, eu, ev, csr, regs, has, ro, ho = _hybrid_setup(p, theta)
    est, secs, st = gpu_hybrid.gpu_hybrid_triangle_count(
        csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v, csr.degrees, regs, has, gamma, p,
        split_kernels=split)
    ref = O.hybrid_count(rp, col, eu, ev, ro, ho, gamma)
    for k in ("n_approx", "n_exact", "n_clamped", "n_fallback"):
        assert st[k] == ref["stats"][k], k
    assert math.isclose(est, ref["total_fsum"], rel_tol=1e-12, abs_tol=1e-9)
    assert math.isclose(est, ref["total_naive"], rel_tol=1e-9)
    if gamma >= 10**9:
        assert st["n_approx"] == 0 and est == float(O.exact_count_csr(rp, col))


@pytest.mark.parametrize("split", [True, False])
def test_hybrid_tuned_knobs_give_the_same_result(device, split):
    # tc_hybrid_count_tuned: the exact part with the exact count's knobs
    # (row trimming, row-ordered items, small chunks, masked tails, no
    # workgroup phase); the supports, hence the estimate and the stats, are
    # those of the default knobs and of the oracle
    p = 10
    rp, col, eu, ev, csr, regs, has, ro, ho = _hybrid_setup(p, 0)
    ref = O.hybrid_count(rp, col, eu, ev, ro, ho, 24)
    base = gpu_hybrid.gpu_hybrid_triangle_count(
        csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v, csr.degrees, regs, has, 24, p,
        split_kernels=split)
    for params in ({"trim_rows": True, "item_order": True, "chunk": 64, "tail_units": True,
                    "block_min_bits": 13, "light_max": 40},
                   {"pack_ids": True, "unroll": 4}):
        est, _, st = gpu_hybrid.gpu_hybrid_triangle_count(
            csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v, csr.degrees, regs, has, 24, p,
            split_kernels=split, params=params)
        assert est == base[0], params
        for k in ("n_approx", "n_exact", "n_clamped", "n_fallback"):
            assert st[k] == base[2][k] == ref["stats"][k], (k, params)
    assert math.isclose(base[0], ref["total_fsum"], rel_tol=1e-12, abs_tol=1e-9)


def test_hybrid_known_answer_all_exact(device):
    # cpu_exp/triangle_hll_experiment.py:1078-1084: threshold 1e9 -> 4, no HLL edges
    csr = gpu_graph.build_csr_gpu(O.SELF_TEST_EDGES)
    regs, has = gpu_hll.alloc_sketch_arrays_device(csr.n_nodes, 8)
    gpu_hll.build_sketches_kernel[1, 256](csr.row_ptr, csr.col_idx, csr.n_nodes, 0, 8, regs, has)
    est, _, st = gpu_hybrid.gpu_hybrid_triangle_count(
        csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v, csr.degrees, regs, has, 10**9, 8)
    assert est == 4.0 and st["n_approx"] == 0


def test_hybrid_per_edge_route_and_determinism(device):
    import torch

    p = 10
    rp, col, eu, ev, csr, regs, has, ro, ho = _hybrid_setup(p, 0)
    rng = np.random.default_rng(3)
    route = (rng.random(len(eu)) < 0.5)
    ref = O.hybrid_count(rp, col, eu, ev, ro, ho, 0, route=route)
    outs = []
    for _ in range(3):
        per = torch.zeros(len(eu), dtype=torch.float64, device="cuda")
        c, s, fb = gpu_hybrid.hybrid_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v,
                                           csr.degrees, regs, has, 0, p,
                                           route=torch.from_numpy(route.astype(np.uint8)),
                                           per_edge=per)
        torch.cuda.synchronize()
        outs.append(float(c.item()))
        np.testing.assert_allclose(per.cpu().numpy(), ref["supports"], rtol=1e-12, atol=1e-12)
        assert int(s[0].item()) == ref["stats"]["n_approx"]
    assert outs[0] == outs[1] == outs[2]
    assert math.isclose(outs[0], ref["total_fsum"], rel_tol=1e-12)


@pytest.mark.parametrize("gamma,theta", [(0, 0), (8, 0), (8, 12), (40, 0), (10**9, 0)])
def test_hybrid_row_length_degrees_and_parts(device, gamma, theta):
    # degrees=None routes and estimates with the row lengths, which are the
    # forward degrees: every output equals the call with the degrees array.
    # The exact part (routed inside the exact pipeline's edge pass) and the HLL
    # part (routed by its own pass) add up to the combined call.
    import torch

    from triangle_counting_hll_amd import _lib

    p = 10
    rp, col, eu, ev, csr, regs, has, ro, ho = _hybrid_setup(p, theta)
    E = len(eu)

    def run(degrees, part=_lib.TC_PART_BOTH):
        per = torch.zeros(E, dtype=torch.float64, device="cuda")
        c, s, fb = gpu_hybrid.hybrid_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v,
                                           degrees, regs, has, gamma, p, part=part, per_edge=per)
        torch.cuda.synchronize()
        return float(c.item()), s.cpu().numpy(), int(fb.item()), per.cpu().numpy()

    c1, s1, f1, p1 = run(csr.degrees)
    c0, s0, f0, p0 = run(None)
    assert c0 == c1 and f0 == f1
    assert np.array_equal(s0, s1) and np.array_equal(p0, p1)
    ref = O.hybrid_count(rp, col, eu, ev, ro, ho, gamma)
    assert int(s1[0]) == ref["stats"]["n_approx"] and f1 == ref["stats"]["n_fallback"]
    ce, se, fe, pe = run(csr.degrees, _lib.TC_PART_EXACT)
    ch, sh, fh, ph = run(csr.degrees, _lib.TC_PART_HLL)
    assert math.isclose(ce + ch, c1, rel_tol=1e-12, abs_tol=1e-9)
    # each part adds its own statistics: exact-routed edges, HLL edges and clamps
    assert se[1] == s1[1] and se[0] == 0 and se[2] == 0
    assert sh[0] == s1[0] and sh[2] == s1[2] and sh[1] == 0
    assert fe == f1 and fh == f1


def test_reference_launch_objects(device):
    import torch

    p = 8
    rp, col, eu, ev, csr, regs, has, ro, ho = _hybrid_setup(p, 0)
    n_edges = len(eu)
    gamma = 6
    ref = O.hybrid_count(rp, col, eu, ev, ro, ho, gamma)
    for mode in ("fused", "split"):
        counts = torch.zeros(1, dtype=torch.float64, device="cuda")
        stats = torch.zeros(3, dtype=torch.int64, device="cuda")
        if mode == "fused":
            gpu_hybrid.hybrid_count_kernel[1, 256](csr.row_ptr, csr.col_idx, csr.edge_u,
                                                   csr.edge_v, n_edges, csr.degrees, regs, has,
                                                   gamma, p, counts, stats)
        else:
            gpu_hybrid.hybrid_exact_part_kernel[1, 256](csr.row_ptr, csr.col_idx, csr.edge_u,
                                                        csr.edge_v, n_edges, csr.degrees, has,
                                                        gamma, counts, stats)
            gpu_hybrid.hybrid_hll_part_kernel[1, 256](csr.edge_u, csr.edge_v, n_edges,
                                                      csr.degrees, regs, has, gamma, p, counts,
                                                      stats)
        torch.cuda.synchronize()
        st = stats.cpu().numpy()
        assert [int(x) for x in st] == [ref["stats"]["n_approx"], ref["stats"]["n_exact"],
                                        ref["stats"]["n_clamped"]]
        assert math.isclose(float(counts.item()), ref["total_fsum"], rel_tol=1e-12)


class _SlotRows:
    """Dense-by-vertex view of compact register rows: ``rows[ids]`` returns
    the rows of those vertices (the oracle indexes registers by vertex id)."""

    def __init__(self, regs, slot):
        self.regs, self.slot = regs, slot

    def __getitem__(self, ids):
        s = self.slot[ids]
        assert (s >= 0).all()
        return self.regs[s]


def test_chunglu_full_size_sketches_and_hybrid(device):
    # the configured 117M-edge power-law graph with the hybrid's p = 12 and
    # sketches for forward degree >= 128 in compact slots (125k vertices):
    # every register equals the C oracle's; per-edge hybrid supports equal the
    # C oracle's exact supports on every exact-routed edge and the numpy
    # oracle's clamped HLL supports on a seeded sample of HLL-routed edges
    import torch

    from oracle import cbind

    p, gamma = 12, 128
    csr = gpu_graph.build_csr_gpu(config_pairs("chunglu"), row_ptr_dtype="int64")
    rp = csr.row_ptr.cpu().numpy()
    col = csr.col_idx.cpu().numpy()
    deg = csr.degrees
    assert np.array_equal(deg.cpu().numpy(), np.diff(rp))
    sk = deg >= gamma
    n_sk = int(sk.sum().item())
    assert n_sk == 125_233
    slot = torch.full_like(deg, -1)
    slot[sk] = torch.arange(n_sk, dtype=torch.int32, device=deg.device)
    regs = torch.zeros(n_sk << p, dtype=torch.uint8, device=deg.device)
    has = torch.zeros_like(deg, dtype=torch.uint8)
    gpu_hll.build_sketches(csr.row_ptr, csr.col_idx, csr.n_nodes, gamma, p, regs, has,
                           slot=slot, accumulate=False)
    slot_h = slot.cpu().numpy().astype(np.int64)
    regs_h = regs.cpu().numpy().reshape(n_sk, 1 << p)
    assert np.array_equal(has.cpu().numpy(), (np.diff(rp) >= gamma).astype(np.uint8))
    assert np.array_equal(regs_h, cbind.hll_registers(rp, col, p, slot_h, n_sk))

    eu, ev = csr.edge_u, csr.edge_v
    per = torch.zeros(csr.n_oriented, dtype=torch.float64, device=deg.device)
    c, s, fb = gpu_hybrid.hybrid_count(csr.row_ptr, csr.col_idx, eu, ev, deg, regs, has, gamma,
                                       p, slot=slot, per_edge=per)
    torch.cuda.synchronize()
    per_h = per.cpu().numpy()
    eu_h = eu.cpu().numpy().astype(np.int64)
    ev_h = ev.cpu().numpy().astype(np.int64)
    d = np.diff(rp)
    use = (np.minimum(d[eu_h], d[ev_h]) >= gamma)
    st = s.cpu().numpy()
    assert int(st[0]) == int(use.sum()) == 30_997_202
    assert int(st[1]) == csr.n_oriented - int(use.sum())
    exact = cbind.edge_supports(rp, col, eu_h, ev_h)
    assert np.array_equal(per_h[~use], exact[~use].astype(np.float64))
    # HLL-routed sample against the numpy oracle (register max, FP64 estimate,
    # support d+u + d+v - union clamped to [0, min])
    idx = np.sort(np.random.default_rng(5).choice(np.nonzero(use)[0], 20_000, replace=False))
    ref = O.hybrid_count(rp, col, eu_h[idx], ev_h[idx], _SlotRows(regs_h, slot_h),
                         has.cpu().numpy(), gamma)
    assert ref["use_hll"].all()
    np.testing.assert_allclose(per_h[idx], ref["supports"], rtol=1e-12, atol=1e-9)
    # the device total is the sum of its per-edge supports (numpy's pairwise
    # sum is within ~1e-15 relative here)
    assert math.isclose(float(c.item()), float(np.sum(per_h)), rel_tol=1e-12)
    # with the exact count's tuned knobs (the bench's hybrid lines): the same
    # per-edge supports, total and stats
    per_t = torch.zeros_like(per)
    tuned = {"light_max": 40, "chunk": 128, "block_min_bits": 11, "trim_rows": True,
             "tail_units": True, "item_order": True}
    c_t, s_t, _ = gpu_hybrid.hybrid_count(csr.row_ptr, csr.col_idx, eu, ev, deg, regs, has,
                                          gamma, p, slot=slot, per_edge=per_t, params=tuned)
    torch.cuda.synchronize()
    assert np.array_equal(per_t.cpu().numpy(), per_h)
    assert float(c_t.item()) == float(c.item())
    assert np.array_equal(s_t.cpu().numpy(), st)
