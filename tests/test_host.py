"""Host-side logic that needs no GPU: generators, batch spec, CSR conversion, launch syntax."""

import numpy as np
import pytest

from oracle import tc_oracle as O
from triangle_counting_hll_amd import batch, gpu_graph, gpu_hll, gpu_hybrid, graphgen


def test_counter_stream_is_splitmix64():
    # key 0, draws i = 0..2 are the published splitmix64 outputs for seed 0
    d = graphgen.draws(0, 0, 3)
    assert [int(x) for x in d] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_rmat_shape_range_and_skew():
    p = graphgen.rmat(12, 8, seed=1)
    assert p.shape == (8 << 12, 2) and p.min() >= 0 and p.max() < (1 << 12)
    # with a = 0.57 the source bit is 0 at a level with probability a + b = 0.76
    top = (p[:, 0] >> 11) & 1
    assert abs(1 - top.mean() - 0.76) < 0.01
    assert np.array_equal(p, graphgen.rmat(12, 8, seed=1))
    assert not np.array_equal(p, graphgen.rmat(12, 8, seed=2))


def test_erdos_renyi_uniform_ends():
    p = graphgen.erdos_renyi(1000, 8 / 1000, seed=0)
    assert len(p) == graphgen.er_pairs_count(1000, 8 / 1000) == 3996
    assert p.min() >= 0 and p.max() < 1000
    assert abs(p.mean() - 499.5) < 15


def test_chung_lu_solver():
    w = graphgen.chung_lu_weights(3_072_441, 2.1, 33000.0, 117e6)
    assert abs(float(w.weight(0)) - 33000.0) < 1e-6
    assert abs(w.alpha - 1 / 1.1) < 1e-12


def test_many_graph_spec():
    s = batch.many_graphs_spec()
    assert s.n_graphs == 4096 and int(s.kind.sum()) == 2048
    er = s.kind == 0
    assert s.param[er].min() >= 200 and s.param[er].max() <= 50000
    assert s.param[~er].min() >= 8 and s.param[~er].max() <= 16
    assert np.array_equal(s.n_pairs[~er], 16 << s.param[~er].astype(np.int64))
    assert np.array_equal(s.n_pairs[er], np.round(4.0 * (s.param[er] - 1)).astype(np.int64))
    assert int(s.id_off[-1]) < 2**31
    s2 = batch.many_graphs_spec()
    assert np.array_equal(s.keys, s2.keys) and np.array_equal(s.param, s2.param)


def test_batch_host_twin_matches_standalone_generators():
    s = batch.many_graphs_spec(12, seed=7, scale_down=5)
    pairs = batch.generate_pairs_numpy(s)
    for g in range(s.n_graphs):
        lo, hi = s.pair_off[g], s.pair_off[g + 1]
        local = pairs[lo:hi] - s.id_off[g]
        if s.kind[g] == 0:
            n = int(s.param[g])
            want = graphgen.erdos_renyi(n, 8.0 / n, seed=7, stream=g)
        else:
            want = graphgen.rmat(int(s.param[g]), 16, seed=7, stream=g)
        assert np.array_equal(local, want)


def test_to_csr_numpy_equals_oracle_layout():
    pairs = graphgen.rmat(9, 8, seed=3)
    edges = {(int(min(a, b)), int(max(a, b))) for a, b in pairs if a != b}
    fwd, _ = O.orient_graph(edges)
    for a, b in zip(gpu_graph.to_csr_numpy(fwd), O.to_csr_numpy(fwd)):
        assert a.dtype == b.dtype and np.array_equal(a, b)
    with pytest.raises(KeyError):
        gpu_graph.to_csr_numpy({0: np.array([5])})


def test_launch_objects_need_subscription():
    with pytest.raises(TypeError):
        gpu_hybrid.exact_count_kernel(1, 2)
    with pytest.raises(TypeError):
        gpu_hll.build_sketches_kernel(1, 2)
    assert callable(gpu_hybrid.hybrid_hll_part_kernel[10, 256])


def test_alloc_sketch_arrays_and_reference_constants():
    regs, has = gpu_hll.alloc_sketch_arrays(10, 6)
    assert regs.shape == (640,) and has.shape == (10,) and not regs.any()
    assert gpu_hll._alpha(16) == 0.673 and gpu_hll._bit_length(5) == 3
    assert gpu_hybrid._THREADS == 256
    gpu_hybrid._require_supported_block_size()
    assert gpu_hybrid.relative_error(0, 0) == 0.0


def test_bench_rmat_parameters_match_named_workloads():
    # bench.py generates the R-MAT workloads with the device twin; its
    # (scale, edge factor, seed) table must name the same graphs
    import bench

    for name, (scale, ef, seed) in bench.RMAT_WORKLOADS.items():
        sd = scale - 9
        assert np.array_equal(graphgen.config_pairs(name, sd), graphgen.rmat(9, ef, seed=seed))


def test_header_constants_match_python_tables():
    import os
    import re

    from triangle_counting_hll_amd import _lib

    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "tchll.h")).read()
    defs = {k: int(v) for k, v in re.findall(r"#define (TC_\w+) (-?\d+)", hdr)}
    assert {"gpu": defs["TC_HLL_GPU"], "cpu_exp": defs["TC_HLL_CPU_EXP"],
            "proto32": defs["TC_HLL_PROTO32"]} == _lib.HASH_MODES
    assert (defs["TC_PART_BOTH"], defs["TC_PART_EXACT"], defs["TC_PART_HLL"]) == (
        _lib.TC_PART_BOTH, _lib.TC_PART_EXACT, _lib.TC_PART_HLL)
    stages = sorted((v, k) for k, v in defs.items() if k.startswith("TC_STAGE_"))
    assert [k[len("TC_STAGE_"):].lower() for _, k in stages] == list(_lib.STAGES)


def test_bench_trimmed_floor_matches_brute_force():
    # the roofline's byte floor with trim_rows: an edge centred on v streams
    # u's ids up to the largest id of v's row; recomputed edge by edge here
    import bench
    from triangle_counting_hll_amd import graphgen

    rp, col, fdeg, _nid, eu, ev, _ = O.csr_from_pairs(graphgen.rmat(10, 8, seed=3))
    light_max, chunk = 8, 16
    got = bench.heavy_floor_bytes(rp, eu, ev, light_max, chunk, col=col, trim=True)
    whole = bench.heavy_floor_bytes(rp, eu, ev, light_max, chunk)
    streamed, centers = 0, []
    for u, v in zip(eu.tolist(), ev.tolist()):
        du, dv = int(fdeg[u]), int(fdeg[v])
        if du == 0 or dv == 0 or du + dv <= light_max:
            continue
        if dv >= du:
            hi = col[rp[v + 1] - 1]
            k = int(np.sum(col[rp[u]:rp[u + 1]] <= hi))
            c = v
        else:
            k, c = dv, u
        if k:
            streamed += 4 * k
            centers.append(c)
    assert got["streamed_rows"] == streamed
    assert got["heavy_edges"] == len(centers) and got["packed_words"] == 8 * len(centers)
    per = np.bincount(np.array(centers, dtype=np.int64), minlength=len(fdeg))
    assert got["center_tables"] == 4 * int((((per + chunk - 1) // chunk) * fdeg).sum())
    assert got["streamed_rows"] < whole["streamed_rows"]
