"""Seeded synthetic graph generators for the benchmark and parity workloads.

The reference only runs on SNAP edge lists read from disk
(``experiment/exp_gpu.py:131-143``); those files are not available here, so
every workload is synthetic.  Each generator returns *raw* undirected pairs as
``int32[m, 2]`` exactly as an edge-list file would hold them: duplicates,
both orientations and self-loops may be present.  Cleaning (canonical
``(min, max)`` pairs, self-loop removal, dedup, dense relabel) is the job of
the CSR builder, as in ``experiment/exp_gpu.py:291-322``.

All randomness comes from ``numpy.random.default_rng(seed)`` (PCG64), so a
workload is fully named by its parameters and seed.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


def _raw_u32_pairs(rng: np.random.Generator, count: int):
    """Yield 32-bit uniform integers, two per PCG64 raw draw."""
    raw = rng.bit_generator.random_raw(count)
    return (raw & np.uint64(0xFFFFFFFF)).astype(np.uint32), (raw >> np.uint64(32)).astype(np.uint32)


def rmat(scale: int, edge_factor: int, a: float = 0.57, b: float = 0.19,
         c: float = 0.19, d: float = 0.05, seed: int = 0) -> np.ndarray:
    """R-MAT pairs: ``edge_factor * 2**scale`` samples over ``2**scale`` ids.

    Each level picks a quadrant with probabilities (a, b, c, d) by comparing a
    32-bit uniform integer against the thresholds ``floor(a * 2**32)``,
    ``floor((a+b) * 2**32)`` and ``floor((a+b+c) * 2**32)``; two levels are
    drawn per 64-bit PCG64 output.  No label permutation is applied.
    """
    if not (1 <= scale <= 30):
        raise ValueError(f"scale must be in [1, 30], got {scale}")
    if abs(a + b + c + d - 1.0) > 1e-9:
        raise ValueError("R-MAT probabilities must sum to 1")
    m = int(edge_factor) << int(scale)
    rng = np.random.default_rng(seed)
    t_a = np.uint32(min(int(a * 2**32), 2**32 - 1))
    t_ab = np.uint32(min(int((a + b) * 2**32), 2**32 - 1))
    t_abc = np.uint32(min(int((a + b + c) * 2**32), 2**32 - 1))
    src = np.zeros(m, dtype=np.int64)
    dst = np.zeros(m, dtype=np.int64)
    level = 0
    while level < scale:
        lo, hi = _raw_u32_pairs(rng, m)
        for r in (lo, hi):
            if level >= scale:
                break
            bit = np.int64(1) << np.int64(scale - 1 - level)
            # quadrant: [0,a) -> (0,0); [a,a+b) -> (0,1); [a+b,a+b+c) -> (1,0); else (1,1)
            s_bit = r >= t_ab
            d_bit = ((r >= t_a) & (r < t_ab)) | (r >= t_abc)
            src |= s_bit.astype(np.int64) * bit
            dst |= d_bit.astype(np.int64) * bit
            level += 1
        del lo, hi
    return np.stack([src, dst], axis=1).astype(np.int32)


def erdos_renyi(n: int, p: float, seed: int = 0) -> np.ndarray:
    """G(n, M) stand-in for G(n, p): ``M = round(p * n * (n - 1) / 2)`` uniform pairs.

    Endpoints are ``(r * n) >> 32`` for 32-bit uniform ``r`` (exact integer
    arithmetic).  Self-loops and repeats are left for the CSR builder to drop.
    """
    if n < 2:
        return np.zeros((0, 2), dtype=np.int32)
    m = int(round(p * n * (n - 1) / 2.0))
    rng = np.random.default_rng(seed)
    lo, hi = _raw_u32_pairs(rng, m)
    u = ((lo.astype(np.uint64) * np.uint64(n)) >> np.uint64(32)).astype(np.int32)
    v = ((hi.astype(np.uint64) * np.uint64(n)) >> np.uint64(32)).astype(np.int32)
    return np.stack([u, v], axis=1)


def rgg_radius(n: int, mean_degree: float) -> float:
    """Radius giving the requested mean degree for n points in the unit square."""
    return math.sqrt(mean_degree / (math.pi * n))


def random_geometric(n: int, mean_degree: float = 2.8, seed: int = 1) -> np.ndarray:
    """Random geometric graph: n uniform points in [0,1)^2 (not a torus).

    Pairs at Euclidean distance ``<= r`` with ``r = sqrt(mean_degree / (pi n))``
    are joined.  Pairs are listed once, ``i < j``.
    """
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    pts = rng.random((n, 2))
    r = rgg_radius(n, mean_degree)
    pairs = cKDTree(pts).query_pairs(r, output_type="ndarray")
    return pairs.astype(np.int32)


@dataclass(frozen=True)
class ChungLuWeights:
    """Continuous power-law weight profile ``w(x) = c (x + i0)^(-alpha)``."""

    n: int
    alpha: float
    c: float
    i0: float

    def weight(self, i):
        return self.c * (np.asarray(i, dtype=np.float64) + self.i0) ** (-self.alpha)


def chung_lu_weights(n: int, beta: float, w_max: float, m_target: float) -> ChungLuWeights:
    """Solve ``w(0) = w_max`` and ``integral_0^n w = 2 m_target`` for (c, i0).

    ``alpha = 1 / (beta - 1)`` gives a degree distribution with exponent beta.
    """
    alpha = 1.0 / (beta - 1.0)
    g = 1.0 - alpha

    def total(i0: float) -> float:
        c = w_max * i0 ** alpha
        if abs(g) < 1e-12:
            return c * (math.log(n + i0) - math.log(i0))
        return c / g * ((n + i0) ** g - i0 ** g)

    lo, hi = 1e-6, 1e9
    for _ in range(200):
        mid = math.sqrt(lo * hi)
        if total(mid) < 2.0 * m_target:
            lo = mid
        else:
            hi = mid
    i0 = math.sqrt(lo * hi)
    return ChungLuWeights(n=n, alpha=alpha, c=w_max * i0 ** alpha, i0=i0)


def chung_lu(n: int, beta: float = 2.1, w_max: float = 33000.0,
             m_target: float = 117e6, oversample: float = 1.02,
             seed: int = 2, chunk: int = 1 << 25) -> np.ndarray:
    """Chung-Lu pairs: ``round(oversample * m_target)`` endpoint pairs drawn
    independently with density proportional to ``w(x)``.

    Endpoints are sampled by the inverse of the continuous CDF of ``w`` on
    ``[0, n)`` and truncated to integers.  Duplicate pairs and self-loops are
    dropped later, which is why the sample is slightly larger than the target.
    """
    wts = chung_lu_weights(n, beta, w_max, m_target)
    g = 1.0 - wts.alpha
    a0 = wts.i0 ** g
    span = (n + wts.i0) ** g - a0
    m = int(round(oversample * m_target))
    rng = np.random.default_rng(seed)
    out = np.empty((m, 2), dtype=np.int32)
    for s in range(0, m, chunk):
        e = min(m, s + chunk)
        u = rng.random((e - s, 2))
        x = (u * span + a0) ** (1.0 / g) - wts.i0
        np.clip(x, 0, n - 1, out=x)
        out[s:e] = x.astype(np.int32)
    return out


# ---------------------------------------------------------------------------
# Named workloads (sizes of the five benchmark configurations)
# ---------------------------------------------------------------------------

def config_pairs(name: str, scale_down: int = 0) -> np.ndarray:
    """Raw pairs for a named workload.

    ``rmat14``   R-MAT scale 14, edge factor 8, (.57,.19,.19,.05), seed 0.
    ``rgg21``    RGG, n = 2**21, mean degree 2.8, seed 1.
    ``chunglu``  Chung-Lu, n = 3,072,441, beta 2.1, w_max 33,000, 117M edges, seed 2.
    ``rmat20``   R-MAT scale 20, edge factor 16, seed 5 (router training graph).
    ``rmat23``   R-MAT scale 23, edge factor 16, seed 4 (router evaluation graph).

    ``scale_down`` shrinks the workload by ``2**scale_down`` for quick tests.
    """
    if name == "rmat14":
        return rmat(14 - scale_down, 8, seed=0)
    if name == "rgg21":
        return random_geometric((1 << 21) >> scale_down, 2.8, seed=1)
    if name == "chunglu":
        f = 1 << scale_down
        return chung_lu(3_072_441 // f, 2.1, 33000.0 / math.sqrt(f), 117e6 / f, seed=2)
    if name == "rmat20":
        return rmat(20 - scale_down, 16, seed=5)
    if name == "rmat23":
        return rmat(23 - scale_down, 16, seed=4)
    raise ValueError(f"unknown workload {name!r}")
