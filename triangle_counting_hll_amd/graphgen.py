"""Seeded synthetic graph generators for the benchmark and parity workloads.

The reference only runs on SNAP edge lists read from disk
(``experiment/exp_gpu.py:131-143``); those files are not available here, so
every workload is synthetic.  Each generator returns *raw* undirected pairs as
``int32[m, 2]`` exactly as an edge-list file would hold them: duplicates,
both orientations and self-loops may be present.  Cleaning (canonical
``(min, max)`` pairs, self-loop removal, dedup, dense relabel) is the job of
the CSR builder, as in ``experiment/exp_gpu.py:291-322``.

R-MAT and Erdos-Renyi pairs come from a counter-based generator (below) so
that the numpy functions here and the HIP kernel behind
``generate_batch_gpu`` produce identical pairs; the random geometric and
Chung-Lu graphs use ``numpy.random.default_rng(seed)`` (PCG64).  A workload is
fully named by its parameters and seed.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# ---------------------------------------------------------------------------
# Counter-based stream: draw(key, i) = mix(key + (i + 1) * GOLDEN), the
# splitmix64 sequence started at state `key`; key(seed, stream) =
# mix(seed * GOLDEN ^ (stream + 1) * STREAM_MUL).  All arithmetic is mod 2^64.
# ---------------------------------------------------------------------------

GOLDEN = 0x9E3779B97F4A7C15
STREAM_MUL = 0xD1B54A32D192ED03
_M64 = (1 << 64) - 1


def _mix_np(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def _mix_int(x: int) -> int:
    x &= _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def stream_key(seed: int, stream: int) -> int:
    """Key of the counter stream `stream` under `seed`."""
    return _mix_int(((seed * GOLDEN) & _M64) ^ (((stream + 1) * STREAM_MUL) & _M64))


def draws(key: int, start: int, count: int) -> np.ndarray:
    """``count`` consecutive 64-bit draws of the stream from index ``start``."""
    i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix_np(np.uint64(key) + i * np.uint64(GOLDEN))


def rmat_thresholds(a: float, b: float, c: float):
    """Quadrant thresholds on a 32-bit uniform: floor(a 2^32), floor((a+b) 2^32), ..."""
    f = lambda x: min(int(x * 2**32), 2**32 - 1)
    return f(a), f(a + b), f(a + b + c)


def _rmat_from_key(key: int, scale: int, m: int, a, b, c) -> np.ndarray:
    t_a, t_ab, t_abc = (np.uint64(t) for t in rmat_thresholds(a, b, c))
    D = (scale + 1) // 2
    src = np.zeros(m, dtype=np.int64)
    dst = np.zeros(m, dtype=np.int64)
    idx = np.arange(m, dtype=np.uint64) * np.uint64(D)
    for j in range(D):
        with np.errstate(over="ignore"):
            r64 = _mix_np(np.uint64(key) + (idx + np.uint64(j + 1)) * np.uint64(GOLDEN))
        for half in range(2):
            level = 2 * j + half
            if level >= scale:
                break
            r = (r64 >> np.uint64(32 * half)) & np.uint64(0xFFFFFFFF)
            bit = np.int64(1) << np.int64(scale - 1 - level)
            s_bit = r >= t_ab
            d_bit = ((r >= t_a) & (r < t_ab)) | (r >= t_abc)
            src |= s_bit.astype(np.int64) * bit
            dst |= d_bit.astype(np.int64) * bit
    return np.stack([src, dst], axis=1).astype(np.int32)


def rmat(scale: int, edge_factor: int, a: float = 0.57, b: float = 0.19,
         c: float = 0.19, d: float = 0.05, seed: int = 0, stream: int = 0) -> np.ndarray:
    """R-MAT pairs: ``edge_factor * 2**scale`` samples over ``2**scale`` ids.

    Pair i uses draws ``i*D .. i*D + D - 1`` of its stream (D = ceil(scale/2));
    level l compares the low (l even) or high (l odd) 32 bits of draw
    ``i*D + l//2`` against the thresholds of ``rmat_thresholds``:
    [0,a) -> (0,0), [a,a+b) -> (0,1), [a+b,a+b+c) -> (1,0), else (1,1), most
    significant level first.  No label permutation is applied.
    """
    if not (1 <= scale <= 30):
        raise ValueError(f"scale must be in [1, 30], got {scale}")
    if abs(a + b + c + d - 1.0) > 1e-9:
        raise ValueError("R-MAT probabilities must sum to 1")
    m = int(edge_factor) << int(scale)
    return _rmat_from_key(stream_key(seed, stream), scale, m, a, b, c)


def er_pairs_count(n: int, p: float) -> int:
    """M = round(p * n * (n - 1) / 2)."""
    return int(round(p * n * (n - 1) / 2.0)) if n >= 2 else 0


def erdos_renyi(n: int, p: float, seed: int = 0, stream: int = 0) -> np.ndarray:
    """G(n, M) stand-in for G(n, p): M = round(p n (n-1) / 2) uniform pairs.

    Pair i takes draw i of its stream: u = (low32 * n) >> 32,
    v = (high32 * n) >> 32 (exact integer arithmetic).  Self-loops and
    repeats are left for the CSR builder to drop.
    """
    m = er_pairs_count(n, p)
    if m == 0:
        return np.zeros((0, 2), dtype=np.int32)
    r = draws(stream_key(seed, stream), 0, m)
    u = ((r & np.uint64(0xFFFFFFFF)) * np.uint64(n)) >> np.uint64(32)
    v = ((r >> np.uint64(32)) * np.uint64(n)) >> np.uint64(32)
    return np.stack([u, v], axis=1).astype(np.int32)


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


def _decimal_width(x: np.ndarray) -> np.ndarray:
    w = np.ones(len(x), dtype=np.int64)
    for k in range(1, 19):
        w += x >= 10 ** k
    return w


def edge_list_text(pairs: np.ndarray, sep: str = "\t", newline: str = "\n",
                   header: str = "") -> bytes:
    """SNAP-style text of ``pairs`` (non-negative ids): one ``u<sep>v<newline>``
    line per pair after ``header``.  Vectorised, for large synthetic inputs."""
    a = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    if len(a) and a.min() < 0:
        raise ValueError("edge_list_text writes non-negative ids only")
    s, nl = sep.encode(), newline.encode()
    wu, wv = _decimal_width(a[:, 0]), _decimal_width(a[:, 1])
    line = wu + len(s) + wv + len(nl)
    start = np.zeros(len(a) + 1, dtype=np.int64)
    np.cumsum(line, out=start[1:])
    buf = np.empty(int(start[-1]), dtype=np.uint8)
    base = start[:-1]

    def digits(col, w, at):
        x = col.copy()
        for k in range(int(w.max(initial=1))):
            m = w > k
            buf[at[m] + w[m] - 1 - k] = (x[m] % 10 + 48).astype(np.uint8)
            x //= 10

    digits(a[:, 0], wu, base)
    for j, ch in enumerate(s):
        buf[base + wu + j] = ch
    digits(a[:, 1], wv, base + wu + len(s))
    for j, ch in enumerate(nl):
        buf[base + wu + len(s) + wv + j] = ch
    return header.encode() + buf.tobytes()
