"""Result rows and JSON schema of the experiment driver against a published
result file of the reference (tests/golden/exp_gpu_schema)."""

import json
import os

from triangle_counting_hll_amd import experiment

GOLD = os.path.join(os.path.dirname(__file__), "golden", "exp_gpu_schema",
                    "ca-hepth_exp_gpu_full.json")


def _gold():
    return json.load(open(GOLD))


def test_row_keys_match_reference_rows():
    gold = _gold()
    by_method = {r["method"]: set(r) for r in gold["rows"]}
    exact = experiment.new_row("GPU Exact (ours)", 5, 0.0, 0, count_time=0.1,
                               count_time_std=0.0, e2e_time=0.1, ref_time=0.1)
    assert set(exact) == by_method["GPU Exact (ours)"]
    dl = experiment.new_row("DOULION q=0.10", 5.0, 0.0, 0.0, count_time=0.1,
                            count_time_std=0.0, ref_time=0.1)
    assert set(dl) == by_method["DOULION q=0.10"]
    hy = experiment.new_row("GPU Hybrid p=8 γ=64", 5.0, 0.0, 0.0, count_time=0.1,
                            count_time_std=0.0, e2e_time=0.2, coverage=0.5, ref_time=0.1,
                            clamp_ratio=0.1, kernel_mode="split_block_reduction",
                            n_sketched=3, build_time=0.01, sketch_alloc_time=0.01)
    assert set(hy) == by_method["GPU Hybrid p=8 γ=64"]


def test_method_sequence_and_sweep_match_reference():
    gold = _gold()
    names = ["GPU Exact (ours)", "cuGraph Exact"]
    names += [f"DOULION q={q:.2f}" for q in experiment.DOULION_QS]
    names += [f"GPU Hybrid p={p} γ={g}" for p in experiment.P_SWEEP
              for g in experiment.GAMMA_SWEEP if not (p == 10 and g == 256)]
    assert names == [r["method"] for r in gold["rows"]]
    assert gold["n_timing_runs"] == experiment.N_TIMING_RUNS
    assert gold["e2e_definition"] == experiment.E2E_DEFINITION


def test_speedups_and_table():
    rows = [experiment.new_row("a", 10, 0.0, 0, count_time=2.0, ref_time=4.0),
            experiment.new_row("b", 12, 0.2, 2, count_time=1.0, e2e_time=4.0, ref_time=4.0)]
    assert rows[0]["spdup_cnt"] == 2.0 and rows[1]["spdup_e2e"] == 1.0
    experiment.fill_cross_speedups(rows, 2.0, None)
    assert rows[1]["spdup_vs_gpu_exact"] == 2.0 and rows[1]["spdup_vs_cugraph"] is None
    text = experiment.format_table(rows, "GPU Exact", 4.0, 2.0, None)
    assert "SUMMARY  |  ref=GPU Exact(4.000s)  GPUExact(2.000s)  cuGraph(N/A)" in text
    assert experiment.bias(12, 10) == 2
