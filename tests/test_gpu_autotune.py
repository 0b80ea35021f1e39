"""The autotuner on the GPU: its timing protocol, and tuned knobs keep counts exact."""

import numpy as np
import pytest

from oracle import cbind
from oracle import tc_oracle as O
from triangle_counting_hll_amd import autotune, batch

pytestmark = pytest.mark.gpu


def test_time_run_is_a_median_of_at_least_five_runs(device):
    calls = []
    ms = autotune.time_run(lambda p: calls.append(p), {"chunk": 512}, reps=2)
    assert len(calls) == 1 + autotune.MIN_REPS and ms >= 0.0


def test_tuned_params_reproduce_oracle_counts(device, tmp_path, monkeypatch):
    monkeypatch.setenv("TCHLL_TUNE_CACHE", str(tmp_path / "tune.json"))
    autotune.clear_memory()
    spec = batch.many_graphs_spec(96, seed=9, scale_down=3)
    pairs = batch.generate_pairs_numpy(spec)
    truth = []
    for g in range(spec.n_graphs):
        rp, col, *_ = O.csr_from_pairs(pairs[spec.pair_off[g]:spec.pair_off[g + 1]])
        truth.append(cbind.exact_count(rp, col))
    truth = np.array(truth, dtype=np.int64)
    b = batch.build_batch(spec)
    seen = []

    def run(params):
        seen.append(dict(params))
        got = batch.batch_exact_count(b, params).cpu().numpy()
        assert np.array_equal(got, truth), params

    params, table = autotune.tune(run, "test-batch", reps=5)
    assert len(table) > 1
    # every knob value of the space was run (and counted exactly) at least once
    for knob, values in autotune.SPACE.items():
        assert {p[knob] for p in seen} >= set(values), knob
    g = batch.BatchCountGraph(b, params)
    assert np.array_equal(g.launch().cpu().numpy(), truth)
    g.close()
    again, table2 = autotune.tune(run, "test-batch")
    assert again == params and table2 == []
    autotune.clear_memory()
