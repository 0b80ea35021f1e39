"""The prototype experiment's HLL variant (hash mode proto32) and its timing
harness on the GPU: registers bit-exact, estimates and per-edge supports within
1e-12 of the oracle, the CSV in the reference's layout."""

import csv
import math
import os

import numpy as np
import pytest

from oracle import tc_oracle as O
from triangle_counting_hll_amd import _lib, gpu_graph, gpu_hll, gpu_hybrid, graphgen, prototype

pytestmark = pytest.mark.gpu


def _setup(pairs, p):
    rp, col, fdeg, nid, eu, ev, _ = O.csr_from_pairs(pairs)
    csr = gpu_graph.build_csr_gpu(pairs)
    regs, has = gpu_hll.alloc_sketch_arrays_device(csr.n_nodes, p)
    gpu_hll.build_sketches(csr.row_ptr, csr.col_idx, csr.n_nodes, 0, p, regs, has,
                           hash_mode="proto32")
    return rp, col, eu, ev, csr, regs, has


@pytest.mark.parametrize("p", [4, 8, 12, 16])
def test_registers_bit_exact(device, p):
    pairs = graphgen.config_pairs("rmat14")
    rp, col, eu, ev, csr, regs, has = _setup(pairs, p)
    assert int(has.sum().item()) == csr.n_nodes  # every vertex sketched at theta 0
    want = O.build_registers_proto(rp, col, p)
    assert np.array_equal(regs.cpu().numpy().reshape(csr.n_nodes, -1), want)


@pytest.mark.parametrize("p", [4, 8, 12])
def test_union_estimates(device, p):
    rp, col, eu, ev, csr, regs, has = _setup(graphgen.rmat(12, 16, seed=9), p)
    R = regs.cpu().numpy().reshape(csr.n_nodes, -1)
    got = gpu_hll.union_estimate_pairs(regs, eu[:3000], ev[:3000], p,
                                       hash_mode="proto32").cpu().numpy()
    want = O.estimate_proto_rows(np.maximum(R[eu[:3000]], R[ev[:3000]]))
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)


@pytest.mark.parametrize("p,gamma", [(8, -1), (8, 4), (8, 32), (4, 16), (12, 8), (16, 2)])
def test_hybrid_per_edge_and_totals(device, p, gamma):
    import torch

    rp, col, eu, ev, csr, regs, has = _setup(graphgen.rmat(12, 16, seed=2), p)
    R = regs.cpu().numpy().reshape(csr.n_nodes, -1)
    want = O.hybrid_count_proto(rp, col, eu, ev, R, gamma)
    per_edge = torch.full((len(eu),), -1.0, dtype=torch.float64, device="cuda")
    counts, stats, fb = gpu_hybrid.hybrid_count(
        csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v, csr.degrees, regs, has, gamma + 1, p,
        per_edge=per_edge, hash_mode="proto32")
    s = stats.cpu().numpy()
    assert int(s[0]) == want["n_heavy"] and int(s[1]) == want["n_light"]
    assert want["n_raised_min"] <= int(s[2]) <= want["n_raised_max"]
    assert int(fb.item()) == 0
    got = per_edge.cpu().numpy()
    light = want["light"]
    np.testing.assert_array_equal(got[light], want["supports"][light])
    np.testing.assert_allclose(got[~light], want["supports"][~light], rtol=1e-12, atol=1e-9)
    assert math.isclose(counts.item(), want["total_fsum"], rel_tol=1e-12, abs_tol=1e-9)


def test_unknown_hash_mode_is_rejected(device):
    import torch

    regs = torch.zeros(2 * 256, dtype=torch.uint8, device="cuda")
    idx = torch.zeros(1, dtype=torch.int32, device="cuda")
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    with pytest.raises(Exception, match="unknown hash mode"):
        _lib.check(_lib.load().tc_union_estimate(_lib.ptr(regs), None, 8, _lib.ptr(idx),
                                                 _lib.ptr(idx), 1, _lib.ptr(out), 3, None))


def test_run_gpu_experiment_writes_reference_table(device, tmp_path):
    pairs = graphgen.rmat(13, 16, seed=11)
    path = tmp_path / "g.txt"
    path.write_bytes(graphgen.edge_list_text(pairs, header="# prototype input\n"))
    N, E, rp, col, deg, eu, ev = prototype.load_and_orient(str(path))
    orp, ocol, ofdeg, onid, oeu, oev, _ = O.csr_from_pairs(pairs)
    assert (N, E) == (len(orp) - 1, len(ocol))
    assert np.array_equal(rp.cpu().numpy(), orp) and np.array_equal(col.cpu().numpy(), ocol)
    gammas = [2, 8, 32]
    out_dir = tmp_path / "results_gpu"
    res = prototype.run_gpu_experiment(N, E, rp, col, deg, eu, ev, gammas, p=8, warmup=1,
                                       repeat=2, out_dir=str(out_dir), verbose=False)
    assert res["exact_count"] == O.exact_count_csr(orp, ocol)
    R = O.build_registers_proto(orp, ocol, 8)
    assert math.isclose(res["hll_estimate"],
                        O.hybrid_count_proto(orp, ocol, oeu, oev, R, -1)["total_fsum"],
                        rel_tol=1e-12)
    for g in gammas:
        want = O.hybrid_count_proto(orp, ocol, oeu, oev, R, g)
        h = res["hybrid_results"][g]
        assert h["n_light"] == want["n_light"] and h["n_heavy"] == want["n_heavy"]
        assert math.isclose(h["estimate"], want["total_fsum"], rel_tol=1e-12)
        assert h["total_time"] > 0 and h["light_time"] > 0 and h["cls_time"] >= 0
    assert res["decision"] in ("GO", "MARGINAL", "NO-GO")
    with open(out_dir / "gpu_timing_results.csv", newline="") as f:
        rows = list(csv.reader(f))
    golden = os.path.join(os.path.dirname(__file__), "golden", "prototype_csv",
                          "gpu_timing_results.csv")
    with open(golden, newline="") as f:
        assert rows[0] == next(csv.reader(f))
    assert [r[0] for r in rows[1:]] == ["all_exact_input", "all_exact_sorted", "all_hll",
                                        "hybrid_g2", "hybrid_g8", "hybrid_g32"]
