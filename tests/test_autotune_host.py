"""The autotuner's search and cache, on a synthetic cost function (no GPU)."""

import json

from triangle_counting_hll_amd import autotune


def _cost(best):
    """Separable cost with its minimum at ``best``; every other value of a knob
    costs 1 + its distance in the value list."""
    def ms(p):
        total = 10.0
        for knob, values in autotune.SPACE.items():
            total += abs(values.index(p[knob]) - values.index(best[knob]))
        return total
    return ms


def test_descent_finds_the_separable_minimum():
    best = {"chunk": 1024, "unroll": 8, "light_max": 16, "waves_per_cu": 8, "block_min_bits": 11,
            "trim_rows": 1, "tail_units": 1, "item_order": 1}
    calls = []

    def measure(p):
        calls.append(dict(p))
        return _cost(best)(p)

    got, ms = autotune.descend(measure, autotune.DEFAULTS, autotune.SPACE)
    assert {k: got[k] for k in best} == best and ms == 10.0
    assert got["plan"] == "sorted"          # knobs outside the space are kept
    # the second sweep confirms the minimum and stops
    assert len(calls) <= 1 + 2 * sum(len(v) for v in autotune.SPACE.values())


def test_descent_ignores_gains_inside_the_margin():
    def measure(p):
        return 100.0 - (0.5 if p["chunk"] == 256 else 0.0) - (5.0 if p["unroll"] == 8 else 0.0)

    got, ms = autotune.descend(measure, autotune.DEFAULTS, autotune.SPACE, margin=0.01)
    assert got["chunk"] == autotune.DEFAULTS["chunk"]   # 0.5% faster: not taken
    assert got["unroll"] == 8 and ms == 95.0            # 5% faster: taken


def test_block_min_bits_is_searched():
    assert "block_min_bits" in autotune.SPACE
    assert autotune.DEFAULTS["block_min_bits"] in autotune.SPACE["block_min_bits"]
    assert all(8 <= b <= 13 for b in autotune.SPACE["block_min_bits"])


def test_cache_round_trip(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    monkeypatch.setenv("TCHLL_TUNE_CACHE", str(path))
    autotune.clear_memory()
    best = {"chunk": 256, "unroll": 16, "light_max": 64, "waves_per_cu": 16, "block_min_bits": 9,
            "trim_rows": 0, "tail_units": 1, "item_order": 1}
    n = []

    def timer(p):
        n.append(1)
        return _cost(best)(p)

    got, table = autotune.tune(None, "wl", timer=timer, device="dev0")
    assert {k: got[k] for k in best} == best and len(table) == len(n) > 0
    # memory hit: nothing is timed again
    got2, table2 = autotune.tune(None, "wl", timer=timer, device="dev0")
    assert got2 == got and table2 == [] and len(n) == len(table)
    # file hit in a fresh process (memory cleared)
    autotune.clear_memory()
    got3, table3 = autotune.tune(None, "wl", timer=timer, device="dev0")
    assert got3 == got and table3 == []
    assert json.loads(path.read_text())["dev0|wl"] == got
    # another device or workload key is searched again
    _, table4 = autotune.tune(None, "wl", timer=timer, device="dev1")
    assert len(table4) > 0
    autotune.clear_memory()


def test_cache_can_be_bypassed(monkeypatch):
    monkeypatch.delenv("TCHLL_TUNE_CACHE", raising=False)
    autotune.clear_memory()
    _, t1 = autotune.tune(None, "k", timer=lambda p: 1.0, device="d")
    _, t2 = autotune.tune(None, "k", timer=lambda p: 1.0, device="d", use_cache=False)
    assert len(t1) > 0 and len(t2) == len(t1)
    autotune.clear_memory()



def test_tune_starts_from_the_given_knobs():
    # the first candidate timed is the starting point with the given knobs
    # over the library defaults; a custom space limits the search to its knobs
    autotune.clear_memory()
    seen = []

    def timer(p):
        seen.append(dict(p))
        return 10.0 + (0.0 if p["light_max"] == 36 else 1.0)

    space = {"light_max": (24, 32, 36, 40)}
    got, table = autotune.tune(None, "wl-start", space=space, timer=timer, device="dev0",
                               use_cache=False, start={"light_max": 40, "chunk": 256})
    assert seen[0] == dict(autotune.DEFAULTS, light_max=40, chunk=256)
    assert got["light_max"] == 36 and got["chunk"] == 256
    assert {p["chunk"] for p, _ in table} == {256}


def test_extra_start_reaches_a_joint_optimum():
    # a setting that pays only when three knobs move together: one knob at a
    # time from the defaults never reaches it, a second descent started there
    # does, and the faster end point wins
    joint = {"item_order": 1, "chunk": 256, "block_min_bits": 11}

    def timer(p):
        if all(p[k] == v for k, v in joint.items()):
            return 90.0
        return 100.0 + sum(1.0 for k, v in joint.items() if p[k] == v)

    autotune.clear_memory()
    got, _ = autotune.tune(None, "joint", timer=timer, device="dev0", use_cache=False)
    assert any(got[k] != v for k, v in joint.items())
    got, table = autotune.tune(None, "joint", timer=timer, device="dev0", use_cache=False,
                               extra_starts=[joint])
    assert {k: got[k] for k in joint} == joint
    keys = [tuple(sorted(p.items())) for p, _ in table]
    assert len(keys) == len(set(keys))          # nothing timed twice
