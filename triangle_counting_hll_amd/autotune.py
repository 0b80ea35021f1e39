"""Autotuner for the exact pipeline's launch knobs.

Knobs (``tc_exact_params``): ``chunk`` (others per work item: the tile of the
intersection kernel), ``unroll`` (streamed ids in flight per lane),
``light_max`` (the merge-versus-hash-probe crossover: edges whose two rows
hold at most this many ids are merged by one lane, the rest are probed by a
wavefront) and ``waves_per_cu`` (resident wavefronts per compute unit).

The search is a coordinate descent from the library defaults, timed on the
caller's own workload with device synchronisation; the winner is cached per
device and workload key in memory and, when ``TCHLL_TUNE_CACHE`` names a file,
in that JSON file.
"""

from __future__ import annotations

import json
import os
import time
from typing import Callable, Dict, Optional

from . import _lib

DEFAULTS = {"chunk": 512, "unroll": 16, "light_max": 32, "waves_per_cu": 16, "plan": "sorted"}
SPACE = {
    "chunk": (256, 512, 1024),
    "unroll": (8, 16),
    "light_max": (16, 32, 64),
    "waves_per_cu": (8, 16),
}

_memory: Dict[str, dict] = {}


def _cache_file() -> Optional[str]:
    return os.environ.get("TCHLL_TUNE_CACHE") or None


def _load_file() -> dict:
    path = _cache_file()
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def _store(key: str, params: dict) -> None:
    _memory[key] = params
    path = _cache_file()
    if path:
        data = _load_file()
        data[key] = params
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def time_run(run: Callable[[dict], None], params: dict, reps: int) -> float:
    """Mean wall milliseconds of ``run(params)`` over ``reps`` synchronised runs."""
    t = _lib.require_device()
    run(params)  # warm
    t.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        run(params)
    t.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def tune(run: Callable[[dict], None], key: str, space: Optional[dict] = None, reps: int = 2,
         use_cache: bool = True, log: Optional[Callable[[str], None]] = None):
    """Best knob setting for ``run`` (coordinate descent); returns ``(params, table)``.

    ``table`` lists every (params, ms) pair that was timed.
    """
    space = space or SPACE
    dev = _lib.device_info()["name"]
    full_key = f"{dev}|{key}"
    if use_cache:
        hit = _memory.get(full_key) or _load_file().get(full_key)
        if hit:
            return dict(hit), []
    best = dict(DEFAULTS)
    table = []
    seen = {}

    def measure(p):
        k = json.dumps(p, sort_keys=True)
        if k not in seen:
            seen[k] = time_run(run, p, reps)
            table.append((dict(p), seen[k]))
            if log:
                log(f"autotune {p} -> {seen[k]:.3f} ms")
        return seen[k]

    best_ms = measure(best)
    for _ in range(2):  # two sweeps over the coordinates
        improved = False
        for knob, values in space.items():
            for v in values:
                cand = dict(best, **{knob: v})
                ms = measure(cand)
                if ms < best_ms * 0.99:
                    best, best_ms, improved = cand, ms, True
        if not improved:
            break
    _store(full_key, best)
    return best, table
