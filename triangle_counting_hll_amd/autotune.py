"""Autotuner for the exact pipeline's launch knobs.

Knobs (``tc_exact_params``): ``chunk`` (others per work item: the tile of the
intersection kernel), ``unroll`` (streamed ids in flight per lane),
``light_max`` (the merge-versus-hash-probe crossover: edges whose two rows
hold at most this many ids are merged in the edge pass, the rest are probed
by a wavefront), ``waves_per_cu`` (resident wavefronts per compute unit) and
``block_min_bits`` (the wavefronts-per-edge choice: centers whose row has at
least ``2^(block_min_bits - 1)`` ids are taken by a whole workgroup -- four
wavefronts over one joined LDS table -- and shorter ones by one wavefront) and
``trim_rows`` (the edge pass cuts each streamed row of u to its ids at most
the center row's largest id: fewer ids streamed where rows are ordered like
the degrees, an extra search per heavy edge everywhere) and ``tail_units``
(each streamed row's last ``len % 4`` ids read with its last whole ids in one
masked 16-byte load, or one by one before its units) and ``item_order`` (the
wavefront phase's work items handed out in order of the position of their
middle streamed row, one range per XCD, or longest center first).

The search is a coordinate descent from the library defaults, each candidate
timed as the median of ``reps`` (at least 5) synchronised runs of the
caller's own workload; a candidate replaces the best one only when it is
faster by more than ``margin``.  The winner is cached per device and workload
key in memory and, when ``TCHLL_TUNE_CACHE`` names a file, in that JSON file.
"""

from __future__ import annotations

import json
import os
import statistics
import time
from typing import Callable, Dict, Optional, Sequence

from . import _lib

DEFAULTS = {"chunk": 512, "unroll": 16, "light_max": 32, "waves_per_cu": 16,
            "block_min_bits": 10, "trim_rows": 0, "tail_units": 0, "item_order": 0,
            "plan": "sorted"}
SPACE = {
    "chunk": (256, 512, 1024),
    "unroll": (8, 16),
    "light_max": (16, 32, 64),
    "waves_per_cu": (8, 16),
    "block_min_bits": (9, 10, 11),
    "trim_rows": (0, 1),
    "tail_units": (0, 1),
    "item_order": (0, 1),
}
MIN_REPS = 5

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


def clear_memory() -> None:
    """Forget the in-process cache (the file cache is left alone)."""
    _memory.clear()


def time_run(run: Callable[[dict], None], params: dict, reps: int = MIN_REPS) -> float:
    """Median wall milliseconds of ``run(params)`` over ``reps`` (>= 5) runs,
    each bracketed by device synchronisations, after one warm-up run."""
    t = _lib.require_device()
    reps = max(int(reps), MIN_REPS)
    run(params)  # warm
    t.cuda.synchronize()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run(params)
        t.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    return float(statistics.median(times))


def descend(measure: Callable[[dict], float], start: dict, space: dict, sweeps: int = 2,
            margin: float = 0.01):
    """Coordinate descent: for each knob in turn, every value is tried with the
    other knobs at the best setting so far; a value wins when it is faster than
    the best by more than ``margin`` (relative).  Stops after ``sweeps`` sweeps
    or a sweep without improvement.  Returns ``(best params, best ms)``."""
    best = dict(start)
    best_ms = measure(best)
    for _ in range(sweeps):
        improved = False
        for knob, values in space.items():
            for v in values:
                cand = dict(best, **{knob: v})
                ms = measure(cand)
                if ms < best_ms * (1.0 - margin):
                    best, best_ms, improved = cand, ms, True
        if not improved:
            break
    return best, best_ms


def tune(run: Callable[[dict], None], key: str, space: Optional[dict] = None,
         reps: int = MIN_REPS, use_cache: bool = True,
         log: Optional[Callable[[str], None]] = None,
         timer: Optional[Callable[[dict], float]] = None, device: Optional[str] = None,
         margin: float = 0.01, start: Optional[dict] = None,
         extra_starts: Sequence[dict] = ()):
    """Best knob setting for ``run`` (coordinate descent); returns ``(params, table)``.

    ``table`` lists every (params, ms) pair that was timed (empty on a cache
    hit).  ``start`` overrides knobs of the starting point (the library
    defaults).  Each of ``extra_starts`` (overrides of that starting point)
    starts one more descent, for settings whose knobs only pay together (one
    knob at a time cannot reach them); the fastest end point wins, and
    settings timed by one descent are not timed again by the next.
    ``timer(params) -> ms`` replaces the device timing of ``run``
    (``time_run``) and ``device`` the device name in the cache key; both exist
    so the search and the cache can be exercised without a GPU.
    """
    space = space or SPACE
    dev = device if device is not None else _lib.device_info()["name"]
    full_key = f"{dev}|{key}"
    if use_cache:
        hit = _memory.get(full_key) or _load_file().get(full_key)
        if hit:
            return dict(hit), []
    table = []
    seen = {}
    timer = timer or (lambda p: time_run(run, p, reps))

    def measure(p):
        k = json.dumps(p, sort_keys=True)
        if k not in seen:
            seen[k] = timer(p)
            table.append((dict(p), seen[k]))
            if log:
                log(f"autotune {p} -> {seen[k]:.3f} ms")
        return seen[k]

    base = dict(DEFAULTS, **(start or {}))
    best, best_ms = descend(measure, base, space, margin=margin)
    for extra in extra_starts:
        cand, ms = descend(measure, dict(base, **extra), space, margin=margin)
        if ms < best_ms:
            best, best_ms = cand, ms
    _store(full_key, best)
    return best, table
