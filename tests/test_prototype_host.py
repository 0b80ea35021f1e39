"""The prototype experiment's CSV writer and verdict against the reference's
published timing table (tests/golden/prototype_csv, written by
experiment1_gpu/gpu_experiment_fixed.py:664-697 on soc-LiveJournal1)."""

import csv
import os

import pytest

from triangle_counting_hll_amd import prototype

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "prototype_csv",
                      "gpu_timing_results.csv")


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.reader(f))


def test_csv_round_trip_of_published_table(tmp_path):
    ref = _rows(GOLDEN)
    assert ref[0] == prototype.CSV_COLUMNS
    by = {r[0]: r for r in ref[1:]}
    # the edge count the table was made on, from its own time x throughput
    E = round(float(by["all_exact_input"][2]) * float(by["all_exact_input"][3]) / 1000.0)
    gammas = [int(r[1]) for r in ref[1:] if r[0].startswith("hybrid_g")]
    hyb = {}
    for g in gammas:
        r = by[f"hybrid_g{g}"]
        hyb[g] = {"total_time": float(r[2]), "light_pct": float(r[5]), "heavy_pct": float(r[6]),
                  "cls_time": float(r[7]), "light_time": float(r[8]), "heavy_time": float(r[9])}
    out = tmp_path / "gpu_timing_results.csv"
    prototype.write_timing_csv(str(out), E, float(by["all_exact_input"][2]),
                               float(by["all_exact_sorted"][2]), float(by["all_hll"][2]),
                               gammas, hyb)
    got = _rows(out)
    assert [r[0] for r in got] == [r[0] for r in ref]
    assert len(got) == len(ref) and all(len(a) == len(b) for a, b in zip(got, ref))
    for a, b in zip(got[1:], ref[1:]):
        for k, (x, y) in enumerate(zip(a, b)):
            if k == 3:  # throughput: the published times are rounded to 1e-4 ms
                assert float(x) == pytest.approx(float(y), rel=2e-6)
            elif k == 4 and a[0] == "all_exact_input":
                assert x == y == "1.00"
            elif k == 4:
                assert float(x) == pytest.approx(float(y), abs=2e-4)
            else:
                assert x == y, (a[0], prototype.CSV_COLUMNS[k], x, y)


def test_verdict_of_published_table():
    ref = {r[0]: r for r in _rows(GOLDEN)[1:]}
    hyb = {int(k[8:]): {"total_time": float(r[2])} for k, r in ref.items()
           if k.startswith("hybrid_g")}
    best, checks, decision = prototype.verdict(float(ref["all_exact_input"][2]),
                                               float(ref["all_exact_sorted"][2]),
                                               float(ref["all_hll"][2]), hyb)
    # the published run: best gamma 256 at 1.02x -- fails the 1.5x check only
    assert best == 256 and checks == [False, True, True] and decision == "MARGINAL"


def test_command_line_defaults():
    a = prototype.parse_args(["--graph", "g.txt"])
    assert (a.gammas, a.p, a.warmup, a.repeat, a.out_dir) == ([32, 64, 128, 256], 8, 3, 5,
                                                              "results_gpu")
    assert prototype.prototype_alpha(256) == pytest.approx(0.7213 / (1 + 1.079 / 256))
    assert prototype.prototype_alpha(16) == 0.673
