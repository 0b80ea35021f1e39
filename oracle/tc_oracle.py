"""CPU oracle for the triangle-counting hot path (test infrastructure only).

This module restates, in numpy and plain Python, the reference algorithms the
MI355X path must match.  It is the *checker*: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``triangle_counting_hll_amd``) never calls it.

Every function cites the reference file:line it restates.  How each function
is pinned (golden vectors, closed forms, library routines, brute force) is
listed in ``tests/test_oracle.py``; nothing here is "parity unpinned" except
the bit-exact *float order* of numpy's pairwise sum inside the estimator,
which the tests bound at 1e-12 relative instead.
"""

from __future__ import annotations

import io
import math
from collections import defaultdict
from typing import Dict, List, Optional, Set, Tuple

import numpy as np

MASK64 = (1 << 64) - 1

# ---------------------------------------------------------------------------
# Hash and HLL register update  (src/hll.py:13-19, 81-96; src/gpu_hll.py:12-17, 47-81)
# ---------------------------------------------------------------------------


def hash64_int(x: int) -> int:
    """splitmix64 finaliser on a Python int (src/hll.py:13-19)."""
    x &= MASK64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK64
    return (x ^ (x >> 31)) & MASK64


def hash64(x: np.ndarray) -> np.ndarray:
    """Vectorised splitmix64 finaliser; uint64 arithmetic wraps mod 2**64."""
    x = np.asarray(x).astype(np.uint64)
    with np.errstate(over="ignore"):
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return x


def hll_index_rank(h: np.ndarray, p: int) -> Tuple[np.ndarray, np.ndarray]:
    """Register index (top p bits) and rank (src/hll.py:82-87).

    ``rho = (64 - p) - bit_length(w) + 1`` for the low ``64 - p`` bits ``w``,
    and ``65 - p`` when ``w == 0``.  bit_length is computed exactly with
    integer comparisons (no float log2).
    """
    h = np.asarray(h, dtype=np.uint64)
    rem = 64 - p
    j = (h >> np.uint64(rem)).astype(np.int64)
    w = h & np.uint64((1 << rem) - 1)
    bl = np.zeros(w.shape, dtype=np.int64)
    for b in range(rem):
        bl += (w >= np.uint64(1 << b)).astype(np.int64)
    rho = rem - bl + 1
    return j, rho.astype(np.uint8)


def hll_add_scalar(registers: np.ndarray, x: int, p: int) -> None:
    """One ``HLL.add`` call on a Python int (src/hll.py:81-91)."""
    h = hash64_int(x)
    rem = 64 - p
    j = h >> rem
    w = h & ((1 << rem) - 1)
    rho = rem + 1 if w == 0 else rem - w.bit_length() + 1
    if rho > registers[j]:
        registers[j] = rho


def build_registers(row_ptr: np.ndarray, col_idx: np.ndarray, p: int,
                    theta: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """Dense register matrix ``uint8[n, 2**p]`` and ``has_sketch uint8[n]``.

    Restates ``build_sketches_kernel`` (src/gpu_hll.py:47-81): node u gets a
    sketch iff ``deg+(u) >= theta``; its registers are the max of the ranks of
    the hashed dense neighbour ids ``col_idx[row_ptr[u]:row_ptr[u+1]]``.
    Same semantics as ``build_sketches`` + ``HLL.add_array``
    (src/hybrid.py:15-41, src/hll.py:93-96) on the dense-id graph.
    """
    if not (4 <= p <= 16):
        raise ValueError(f"p must be in [4, 16], got {p}")
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    n = len(row_ptr) - 1
    m = 1 << p
    deg = np.diff(row_ptr)
    has = (deg >= theta).astype(np.uint8)
    regs = np.zeros((n, m), dtype=np.uint8)
    owner = np.repeat(np.arange(n, dtype=np.int64), deg)
    keep = has[owner].astype(bool)
    j, rho = hll_index_rank(hash64(np.asarray(col_idx, dtype=np.int64)[keep]), p)
    flat = owner[keep] * m + j
    out = regs.reshape(-1)
    np.maximum.at(out, flat, rho)
    return regs, has


# ---------------------------------------------------------------------------
# Estimator  (src/hll.py:26-54, 112-120)
# ---------------------------------------------------------------------------


def alpha(m: int) -> float:
    """Bias-correction constant (src/hll.py:26-34)."""
    if m == 16:
        return 0.673
    if m == 32:
        return 0.697
    if m == 64:
        return 0.709
    return 0.7213 / (1.0 + 1.079 / m)


def estimate_from_registers(registers: np.ndarray, alpha_m: Optional[float] = None) -> float:
    """FP64 HLL estimate of one register row (src/hll.py:37-54)."""
    registers = np.asarray(registers)
    m = len(registers)
    if alpha_m is None:
        alpha_m = alpha(m)
    Z = float(np.sum(np.power(2.0, -registers.astype(np.float64))))
    E = alpha_m * m * m / Z
    if E <= 2.5 * m:
        V = int(np.sum(registers == 0))
        if V > 0:
            E = m * np.log(m / V)
    two64 = float(1 << 64)
    if E > two64 / 30.0:
        E = -two64 * np.log(1.0 - E / two64)
    return float(E)


def estimate_rows(regs: np.ndarray) -> np.ndarray:
    """``estimate_from_registers`` applied to each row of ``uint8[k, m]``."""
    regs = np.asarray(regs)
    k, m = regs.shape
    a = alpha(m)
    Z = np.sum(np.power(2.0, -regs.astype(np.float64)), axis=1)
    E = a * m * m / Z
    V = np.sum(regs == 0, axis=1)
    small = (E <= 2.5 * m) & (V > 0)
    with np.errstate(divide="ignore"):
        E = np.where(small, m * np.log(m / np.maximum(V, 1)), E)
    two64 = float(1 << 64)
    big = E > two64 / 30.0
    if np.any(big):
        with np.errstate(invalid="ignore", divide="ignore"):
            E = np.where(big, -two64 * np.log(1.0 - E / two64), E)
    return E


def union_estimate_rows(regs_a: np.ndarray, regs_b: np.ndarray) -> np.ndarray:
    """Register-wise max then estimate, row by row (src/hll.py:112-120)."""
    return estimate_rows(np.maximum(regs_a, regs_b))


# ---------------------------------------------------------------------------
# Orientation and CSR  (src/graph.py:29-61; src/gpu_graph.py:6-41;
# experiment/exp_gpu.py:291-365)
# ---------------------------------------------------------------------------


def orient_graph(edges: Set[Tuple[int, int]]) -> Tuple[Dict[int, np.ndarray], Dict[int, int]]:
    """Dict-based orientation by (degree, id) (src/graph.py:29-61).  Small graphs only."""
    degree: Dict[int, int] = defaultdict(int)
    for u, v in edges:
        degree[u] += 1
        degree[v] += 1
    fwd: Dict[int, list] = defaultdict(list)
    nodes: Set[int] = set()
    for u, v in edges:
        nodes.add(u)
        nodes.add(v)
        if (degree[u], u) < (degree[v], v):
            fwd[u].append(v)
        else:
            fwd[v].append(u)
    return {x: np.array(sorted(fwd[x]), dtype=np.int64) for x in nodes}, dict(degree)


def to_csr_numpy(forward_adj: Dict[int, np.ndarray]):
    """Dense-relabelled int32 CSR from a forward-adjacency dict (src/gpu_graph.py:6-41)."""
    node_ids = np.array(sorted(forward_adj.keys()), dtype=np.int32)
    n = len(node_ids)
    index = {int(x): i for i, x in enumerate(node_ids)}
    degrees = np.array([len(forward_adj[int(x)]) for x in node_ids], dtype=np.int32)
    row_ptr = np.zeros(n + 1, dtype=np.int32)
    row_ptr[1:] = np.cumsum(degrees)
    E = int(row_ptr[-1])
    col_idx = np.empty(E, dtype=np.int32)
    edge_u = np.empty(E, dtype=np.int32)
    edge_v = np.empty(E, dtype=np.int32)
    for i, x in enumerate(node_ids):
        s = row_ptr[i]
        for k, v in enumerate(forward_adj[int(x)]):
            j = index[int(v)]
            col_idx[s + k] = j
            edge_u[s + k] = i
            edge_v[s + k] = j
    return row_ptr, col_idx, degrees, node_ids, edge_u, edge_v


# ---------------------------------------------------------------------------
# The CPU experiment's HLL variant  (cpu_exp/triangle_hll_experiment.py:45-48,
# 347-398, 615-709)
# ---------------------------------------------------------------------------

SM64_GOLDEN = 0x9E3779B97F4A7C15


def splitmix64_seeded_int(x: int) -> int:
    """splitmix64 with the golden increment (cpu_exp:369-373) on a Python int."""
    return hash64_int((x + SM64_GOLDEN) & MASK64)


def rho_high_int(h: int, p: int) -> int:
    """Rank from the high 64 - p bits (cpu_exp:376-381)."""
    bits = 64 - p
    w = h >> p
    return bits + 1 if w == 0 else bits - w.bit_length() + 1


def build_registers_cpu_exp_py(row_ptr, col_idx, p: int, seed: int) -> np.ndarray:
    """``build_hll_registers_py`` (cpu_exp:384-398): uint8 [n, 2^p], one Python
    step per element (small graphs only)."""
    n = len(row_ptr) - 1
    m = 1 << p
    regs = np.zeros((n, m), dtype=np.uint8)
    seed &= MASK64
    for u in range(n):
        for k in range(int(row_ptr[u]), int(row_ptr[u + 1])):
            h = splitmix64_seeded_int((int(col_idx[k]) + seed) & MASK64)
            j = h & (m - 1)
            r = rho_high_int(h, p)
            if r > regs[u, j]:
                regs[u, j] = r
    return regs


def build_registers_cpu_exp(row_ptr, col_idx, p: int, seed: int) -> np.ndarray:
    """Vectorised ``build_registers_cpu_exp_py``: uint8 [n, 2^p]."""
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    n = len(row_ptr) - 1
    m = 1 << p
    x = np.asarray(col_idx).astype(np.uint64)
    with np.errstate(over="ignore"):
        h = hash64(x + np.uint64(seed & MASK64) + np.uint64(SM64_GOLDEN))
    j = (h & np.uint64(m - 1)).astype(np.int64)
    w = h >> np.uint64(p)
    bl = np.zeros(w.shape, dtype=np.int64)
    for b in range(64 - p):
        bl += (w >= np.uint64(1 << b)).astype(np.int64)
    rho = (64 - p) - bl + 1  # w == 0 gives bl = 0 and rho = 65 - p
    u = np.repeat(np.arange(n, dtype=np.int64), np.diff(row_ptr))
    regs = np.zeros(n * m, dtype=np.uint8)
    np.maximum.at(regs, u * m + j, rho.astype(np.uint8))
    return regs.reshape(n, m)


def estimate_cpu_exp(registers: np.ndarray) -> float:
    """``estimate_hll_cardinality`` (cpu_exp:356-365)."""
    r = np.asarray(registers)
    m = int(r.size)
    inv_sum = float(np.sum(np.exp2(-r.astype(np.float64))))
    raw = alpha(m) * m * m / inv_sum
    zeros = int(np.count_nonzero(r == 0))
    if raw <= 2.5 * m and zeros > 0:
        return m * math.log(m / zeros)
    if raw > 4294967296.0 / 30.0:
        return -4294967296.0 * math.log(1.0 - raw / 4294967296.0)
    return raw


def union_estimate_cpu_exp(ra: np.ndarray, rb: np.ndarray) -> float:
    """``hll_union_estimate_numba`` (cpu_exp:615-635): register-wise max, then
    2^-r summed left to right."""
    r = np.maximum(np.asarray(ra), np.asarray(rb)).astype(np.int64)
    m = int(r.size)
    inv_sum = float(np.add.accumulate(np.ldexp(1.0, -r))[-1])
    zeros = int(np.count_nonzero(r == 0))
    raw = alpha(m) * m * m / inv_sum
    if raw <= 2.5 * m and zeros > 0:
        return m * math.log(m / zeros)
    if raw > 4294967296.0 / 30.0:
        return -4294967296.0 * math.log(1.0 - raw / 4294967296.0)
    return raw


def hll_support_cpu_exp(regs, out_degree, u: int, v: int) -> Tuple[float, bool]:
    """(clamped support, clamped?) of one edge (cpu_exp:637-660)."""
    est = float(out_degree[u] + out_degree[v]) - union_estimate_cpu_exp(regs[u], regs[v])
    upper = float(min(out_degree[u], out_degree[v]))
    if est < 0.0:
        return 0.0, True
    if est > upper:
        return upper, True
    return est, False


def pure_hll_count_cpu_exp(edge_u, edge_v, out_degree, regs) -> Tuple[float, int]:
    """``pure_hll_count`` (cpu_exp:662-684): (total, clamp_count)."""
    total, clamps = [], 0
    for u, v in zip(np.asarray(edge_u).tolist(), np.asarray(edge_v).tolist()):
        s, c = hll_support_cpu_exp(regs, out_degree, u, v)
        total.append(s)
        clamps += c
    return math.fsum(total), clamps


def hybrid_count_cpu_exp(row_ptr, col_idx, edge_u, edge_v, out_degree, regs, threshold: int):
    """``hybrid_count`` (cpu_exp:686-709): exact when min(out deg) <= threshold.
    Returns (total, exact_edges, hll_edges, exact_sum, hll_sum, clamp_count)."""
    eu = np.asarray(edge_u)
    ev = np.asarray(edge_v)
    cost = np.minimum(out_degree[eu], out_degree[ev])
    ex = cost <= threshold
    exact_sum = float(edge_supports(row_ptr, col_idx, eu[ex], ev[ex]).sum())
    hll_sum, clamps = pure_hll_count_cpu_exp(eu[~ex], ev[~ex], out_degree, regs)
    return (exact_sum + hll_sum, int(ex.sum()), int((~ex).sum()), exact_sum, hll_sum, clamps)


# ---------------------------------------------------------------------------
# The prototype GPU experiment's HLL variant  (experiment1_gpu/
# gpu_experiment_fixed.py:131-166, 209-254, 257-271, 397-417, 437-526;
# experiment1_gpu/gpu_experiment.py:137-175 builds the same registers on the CPU)
# ---------------------------------------------------------------------------

MASK32 = (1 << 32) - 1
PROTO_SEED32 = 0x9747B28C
PROTO_MULT32 = 0x45D9F3B


def proto_hash32_int(v: int) -> int:
    """The prototype's 32-bit mixer of an id (gpu_experiment_fixed.py:143-149)."""
    h = PROTO_SEED32 ^ (v & MASK32)
    h ^= h >> 16
    h = (h * PROTO_MULT32) & MASK32
    h ^= h >> 16
    h = (h * PROTO_MULT32) & MASK32
    return h ^ (h >> 16)


def proto_rank_loop(h: int, p: int) -> int:
    """Rank by the prototype's bit loop (gpu_experiment_fixed.py:151-158): one
    plus the zeros above the first set bit of ``(h << p) | 1`` scanned from
    bit 31 - p down to bit 0."""
    remaining = ((h << p) | 1) & MASK32
    rho = 1
    for bit in range(31 - p, -1, -1):
        if remaining & (1 << bit):
            break
        rho += 1
    return rho


def proto_hash32(v: np.ndarray) -> np.ndarray:
    """``proto_hash32_int`` over a uint32 array."""
    h = np.asarray(v).astype(np.uint32) ^ np.uint32(PROTO_SEED32)
    with np.errstate(over="ignore"):
        h = h ^ (h >> np.uint32(16))
        h = h * np.uint32(PROTO_MULT32)
        h = h ^ (h >> np.uint32(16))
        h = h * np.uint32(PROTO_MULT32)
    return h ^ (h >> np.uint32(16))


def proto_index_rank(h: np.ndarray, p: int) -> Tuple[np.ndarray, np.ndarray]:
    """Bucket (top p bits) and rank (closed form of ``proto_rank_loop``: the
    scanned window is the low 32 - p bits of ``(h << p) | 1``, never zero)."""
    h = np.asarray(h).astype(np.uint64)
    j = (h >> np.uint64(32 - p)).astype(np.int64)
    window = ((h << np.uint64(p)) | np.uint64(1)) & np.uint64((1 << (32 - p)) - 1)
    bl = np.zeros(window.shape, dtype=np.int64)
    for b in range(32 - p):
        bl += (window >= np.uint64(1 << b)).astype(np.int64)
    return j, (32 - p) - bl + 1


def build_registers_proto_py(row_ptr, col_idx, p: int) -> np.ndarray:
    """``build_sketches`` (gpu_experiment_fixed.py:131-166) one element at a
    time: uint8 [n, 2^p], every vertex sketched (small graphs only)."""
    n = len(row_ptr) - 1
    m = 1 << p
    regs = np.zeros((n, m), dtype=np.uint8)
    for u in range(n):
        for k in range(int(row_ptr[u]), int(row_ptr[u + 1])):
            h = proto_hash32_int(int(col_idx[k]))
            j = (h >> (32 - p)) & (m - 1)
            r = proto_rank_loop(h, p)
            if r > regs[u, j]:
                regs[u, j] = r
    return regs


def build_registers_proto(row_ptr, col_idx, p: int) -> np.ndarray:
    """Vectorised ``build_registers_proto_py``: uint8 [n, 2^p]."""
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    n = len(row_ptr) - 1
    m = 1 << p
    j, rho = proto_index_rank(proto_hash32(np.asarray(col_idx)), p)
    u = np.repeat(np.arange(n, dtype=np.int64), np.diff(row_ptr))
    regs = np.zeros(n * m, dtype=np.uint8)
    np.maximum.at(regs, u * m + j, rho.astype(np.uint8))
    return regs.reshape(n, m)


def estimate_proto_rows(regs: np.ndarray) -> np.ndarray:
    """The prototype's estimator per row (gpu_experiment_fixed.py:236-249) in
    FP64: raw ``alpha m^2 / Z``, LinearCounting when raw <= 2.5 m and a
    register is zero, no large-range correction."""
    regs = np.asarray(regs)
    k, m = regs.shape
    Z = np.sum(np.ldexp(1.0, -regs.astype(np.int64)), axis=1)
    E = alpha(m) * m * m / Z
    V = np.sum(regs == 0, axis=1)
    small = (E <= 2.5 * m) & (V > 0)
    with np.errstate(divide="ignore"):
        return np.where(small, m * np.log(m / np.maximum(V, 1)), E)


def support_proto_rows(ra: np.ndarray, rb: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """``hll_union`` (gpu_experiment_fixed.py:209-254) per row pair in FP64:
    ``est(a) + est(b) - est(max(a, b))`` raised to 0 when negative.
    Returns (supports, raised-to-zero flags)."""
    s = estimate_proto_rows(ra) + estimate_proto_rows(rb) - estimate_proto_rows(np.maximum(ra, rb))
    neg = s < 0.0
    return np.where(neg, 0.0, s), neg


def support_proto_fp32(ra: np.ndarray, rb: np.ndarray) -> float:
    """The reference kernel's own FP32 arithmetic for one edge
    (gpu_experiment_fixed.py:224-253): float sums in register order, float
    alpha, logf.  Used to bound the FP64 restatement's distance from it."""
    m = len(ra)
    f = np.float32
    a = f(0.7213 / (1.0 + 1.079 / m)) if m >= 128 else f({16: 0.673, 32: 0.697, 64: 0.709}[m])
    su = sv = sw = f(0.0)
    zu = zv = zw = 0
    for k in range(m):
        x, y = int(ra[k]), int(rb[k])
        w = max(x, y)
        su = f(su + np.exp2(f(-x)))
        sv = f(sv + np.exp2(f(-y)))
        sw = f(sw + np.exp2(f(-w)))
        zu += x == 0
        zv += y == 0
        zw += w == 0
    fm = f(m)
    est = []
    for s_, z in ((su, zu), (sv, zv), (sw, zw)):
        e = f(a * fm * fm / s_)
        if e <= f(2.5) * fm and z > 0:
            e = f(fm * np.log(f(fm / f(z))))
        est.append(e)
    s = f(est[0] + est[1] - est[2])
    return float(s) if s > 0 else 0.0


def hybrid_count_proto(row_ptr, col_idx, edge_u, edge_v, regs: np.ndarray, gamma: int) -> dict:
    """The prototype's hybrid (gpu_experiment_fixed.py:257-271, 437-526):
    light iff ``min(deg+u, deg+v) <= gamma`` (exact merge count), heavy edges
    by ``support_proto_rows``.  The prototype times the phases and keeps the
    per-edge values on the device; this returns them with the totals."""
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    eu = np.asarray(edge_u, dtype=np.int64)
    ev = np.asarray(edge_v, dtype=np.int64)
    deg = np.diff(row_ptr)
    light = np.minimum(deg[eu], deg[ev]) <= gamma
    sup = np.zeros(len(eu))
    sup[light] = edge_supports(row_ptr, col_idx, eu[light], ev[light])
    heavy = np.nonzero(~light)[0]
    raw = np.zeros(len(eu))
    scale = np.zeros(len(eu))
    n_neg = 0
    if len(heavy):
        est = estimate_proto_rows(regs)  # per vertex, reused by every edge
        ra, rb = regs[eu[heavy]], regs[ev[heavy]]
        ea, eb = est[eu[heavy]], est[ev[heavy]]
        raw[heavy] = ea + eb - estimate_proto_rows(np.maximum(ra, rb))
        neg = raw[heavy] < 0.0
        sup[heavy] = np.where(neg, 0.0, raw[heavy])
        n_neg = int(neg.sum())
        scale[heavy] = np.maximum(ea, eb)
    exact_part = int(np.sum(sup[light]))
    hll_part = math.fsum(sup[heavy].tolist())
    # raw supports that are 0 in exact arithmetic (e.g. LinearCounting with
    # m * V(u|v) == V(u) * V(v)) come out as +-1e-13: their raise-to-0 flag
    # is a rounding tie, bounded by n_raised_min .. n_raised_max.  An exact 0
    # (an empty row: est 0 and est(u | v) == est(u) bit for bit) is no tie.
    tie = (np.abs(raw) <= 1e-9 * scale) & (raw != 0)
    return {"supports": sup, "light": light, "n_light": int(light.sum()),
            "n_heavy": int(len(heavy)), "n_raised": n_neg, "s_hat": raw,
            "n_raised_min": int(((raw < 0) & ~tie & ~light).sum()),
            "n_raised_max": int(((raw < 0) | tie)[~light].sum()),
            "exact_part": exact_part, "hll_part_fsum": hll_part,
            "total_fsum": exact_part + hll_part}


# ---------------------------------------------------------------------------
# Edge-list text  (experiment/exp_gpu.py:93, 131-143, 291-313)
# ---------------------------------------------------------------------------

COMMENT_CHARS = "#%"  # experiment/exp_gpu.py:93


def iter_edges_raw(text: bytes, comment_chars: str = COMMENT_CHARS):
    """``_iter_edges_raw`` (experiment/exp_gpu.py:131-143) over an in-memory file.

    The text is read as the reference's ``open(path)`` reads a file: UTF-8 with
    universal newlines, one ``line`` per iteration.
    """
    with io.TextIOWrapper(io.BytesIO(text), encoding="utf-8", newline=None) as f:
        for line in f:
            line = line.strip()
            if not line or line[0] in comment_chars:
                continue
            parts = line.split()
            if len(parts) < 2:
                continue
            u, v = int(parts[0]), int(parts[1])
            if u != v:
                yield u, v


def parse_edge_list(text: bytes, comment_chars: str = COMMENT_CHARS) -> np.ndarray:
    """Raw pairs int64 [m, 2] in file order (the edge buffer of exp_gpu.py:299-313,
    before canonicalisation).  Ids must fit the reference's int32 buffer
    (exp_gpu.py:300): negative ids raise ValueError, larger ones OverflowError."""
    out = []
    for u, v in iter_edges_raw(text, comment_chars):
        if u < 0 or v < 0:
            raise ValueError("negative node id")
        if u > 2**31 - 1 or v > 2**31 - 1:
            raise OverflowError("node id does not fit int32")
        out.append((u, v))
    return np.array(out, dtype=np.int64).reshape(-1, 2)


def csr_from_pairs(pairs: np.ndarray):
    """Vectorised loader restating ``_load_graph`` (experiment/exp_gpu.py:291-365).

    Raw pairs -> drop self-loops -> canonical (min, max) -> unique -> dense
    relabel by sorted unique endpoint -> undirected degree -> orient by
    (degree, id) -> forward CSR with rows sorted by (src, dst).

    Returns ``(row_ptr int64[n+1], col_idx int32[E], fwd_deg int32[n],
    node_ids int64[n], edge_u int32[E], edge_v int32[E], undirected_deg int32[n])``.
    """
    pairs = np.asarray(pairs)
    a = np.minimum(pairs[:, 0], pairs[:, 1]).astype(np.int64)
    b = np.maximum(pairs[:, 0], pairs[:, 1]).astype(np.int64)
    keep = a != b
    key = np.unique((a[keep] << 32) | b[keep])
    a = key >> 32
    b = key & 0xFFFFFFFF
    if len(key) == 0:
        z32 = np.zeros(0, dtype=np.int32)
        return (np.zeros(1, dtype=np.int64), z32, z32, np.zeros(0, dtype=np.int64), z32, z32, z32)
    node_ids = np.union1d(a, b)
    n = len(node_ids)
    nidx = np.full(int(node_ids[-1]) + 1, -1, dtype=np.int64)
    nidx[node_ids] = np.arange(n)
    ra = nidx[a]
    rb = nidx[b]
    deg = (np.bincount(ra, minlength=n) + np.bincount(rb, minlength=n)).astype(np.int64)
    first = (deg[ra] < deg[rb]) | ((deg[ra] == deg[rb]) & (ra < rb))
    src = np.where(first, ra, rb)
    dst = np.where(first, rb, ra)
    fdeg = np.bincount(src, minlength=n)
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    row_ptr[1:] = np.cumsum(fdeg)
    order = np.lexsort((dst, src))
    col = dst[order].astype(np.int32)
    eu = src[order].astype(np.int32)
    return (row_ptr, col, fdeg.astype(np.int32), node_ids, eu, col.copy(), deg.astype(np.int32))


def forward_adj_from_csr(row_ptr, col_idx) -> Dict[int, np.ndarray]:
    """``forward_adj`` dict over dense ids (experiment/exp_gpu.py:368-377)."""
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    return {i: np.asarray(col_idx[row_ptr[i]:row_ptr[i + 1]], dtype=np.int64)
            for i in range(len(row_ptr) - 1)}


# ---------------------------------------------------------------------------
# Exact counting  (src/exact.py:12-98)
# ---------------------------------------------------------------------------


def intersect_adaptive(a: np.ndarray, b: np.ndarray, r: float = 8.0) -> int:
    """Merge when size ratio <= r, else binary search (src/exact.py:29-41)."""
    la, lb = len(a), len(b)
    if la == 0 or lb == 0:
        return 0
    if la > lb:
        a, b, la, lb = b, a, lb, la
    if lb / la <= r:
        return int(len(np.intersect1d(a, b, assume_unique=True)))
    idx = np.searchsorted(b, a)
    mask = idx < lb
    return int(np.sum(b[idx[mask]] == a[mask]))


def exact_triangle_count(forward_adj: Dict[int, np.ndarray], r: float = 8.0) -> int:
    """Forward-exact count over the dict (src/exact.py:48-98).  Small graphs."""
    total = 0
    for u, nu in forward_adj.items():
        if len(nu) == 0:
            continue
        for v in nu:
            nv = forward_adj.get(int(v))
            if nv is None or len(nv) == 0:
                continue
            total += intersect_adaptive(nu, nv, r)
    return total


def edge_supports(row_ptr, col_idx, edge_u, edge_v) -> np.ndarray:
    """Per-edge ``|N+(u) ∩ N+(v)|`` for the given edge list (src/exact.py:79-95).

    Vectorised: every (edge, element of N+(u)) pair is tested for membership in
    N+(v) by ``searchsorted`` on the globally sorted keys ``row * 2**32 + col``.
    Memory is O(sum over edges of deg+(u)); use for graphs up to ~1e7 of that.
    """
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    col = np.asarray(col_idx, dtype=np.int64)
    eu = np.asarray(edge_u, dtype=np.int64)
    ev = np.asarray(edge_v, dtype=np.int64)
    n = len(row_ptr) - 1
    deg = np.diff(row_ptr)
    owner = np.repeat(np.arange(n, dtype=np.int64), deg)
    keys = (owner << 32) | col            # sorted because rows are sorted
    du = deg[eu]
    tot = int(du.sum())
    out = np.zeros(len(eu), dtype=np.int64)
    if tot == 0:
        return out
    e_rep = np.repeat(np.arange(len(eu), dtype=np.int64), du)
    starts = np.repeat(row_ptr[eu], du)
    offs = np.arange(tot, dtype=np.int64) - np.repeat(np.cumsum(du) - du, du)
    w = col[starts + offs]
    probe = (ev[e_rep] << 32) | w
    pos = np.searchsorted(keys, probe)
    hit = np.zeros(tot, dtype=bool)
    ok = pos < len(keys)
    hit[ok] = keys[pos[ok]] == probe[ok]
    np.add.at(out, e_rep[hit], 1)
    return out


def exact_count_csr(row_ptr, col_idx, edge_u=None, edge_v=None) -> int:
    """Exact count over the CSR's COO edge list (same sum as src/exact.py:79-95)."""
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    if edge_u is None:
        deg = np.diff(row_ptr)
        edge_u = np.repeat(np.arange(len(deg)), deg)
        edge_v = col_idx
    return int(edge_supports(row_ptr, col_idx, edge_u, edge_v).sum())


# ---------------------------------------------------------------------------
# Hybrid  (src/hybrid.py:48-166; src/gpu_hybrid.py:80-89, 137-211, 357-431)
# ---------------------------------------------------------------------------


def hybrid_count(row_ptr, col_idx, edge_u, edge_v, regs: np.ndarray, has: np.ndarray,
                 gamma: int, route: Optional[np.ndarray] = None) -> dict:
    """Per-edge hybrid estimate on dense arrays.

    Routing (src/hybrid.py:107-112, src/gpu_hybrid.py:80-89): HLL iff
    ``min(deg+u, deg+v) >= gamma`` and both endpoints have sketches; a given
    boolean ``route`` array (True = wants HLL) replaces the degree rule.  HLL
    support ``deg+u + deg+v - union_estimate`` clamped to ``[0, min]``
    (src/hybrid.py:117-125).  Returns per-edge supports, the naive running
    total (src/hybrid.py:125/134 order), an ``fsum`` total, and the stats keys
    of src/hybrid.py:152-164.
    """
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    eu = np.asarray(edge_u, dtype=np.int64)
    ev = np.asarray(edge_v, dtype=np.int64)
    deg = np.diff(row_ptr)
    du = deg[eu]
    dv = deg[ev]
    proxy = np.minimum(du, dv)
    wants = proxy >= gamma if route is None else np.asarray(route, dtype=bool)
    both = (has[eu] == 1) & (has[ev] == 1)
    use = wants & both
    sup = edge_supports(row_ptr, col_idx, eu, ev).astype(np.float64)
    s_hat = np.zeros(len(eu))
    clamped = np.zeros(len(eu), dtype=bool)
    idx = np.nonzero(use)[0]
    if len(idx):
        est = union_estimate_rows(regs[eu[idx]], regs[ev[idx]])
        s = du[idx] + dv[idx] - est
        up = proxy[idx].astype(np.float64)
        clamped[idx] = (s < 0.0) | (s > up)
        s_hat[idx] = s
        sup[idx] = np.maximum(0.0, np.minimum(s, up))
    n_edges = len(eu)
    n_approx = int(use.sum())
    n_fallback = int((wants & ~both).sum())
    n_clamped = int(clamped.sum())
    naive = 0.0
    for x in sup.tolist():
        naive += x
    return {
        "supports": sup,
        "use_hll": use,
        "s_hat": s_hat,
        "total_naive": naive,
        "total_fsum": math.fsum(sup.tolist()),
        "exact_part": int(np.sum(sup[~use])),
        "hll_part_fsum": math.fsum(sup[use].tolist()),
        "stats": {
            "n_edges": n_edges,
            "n_approx": n_approx,
            "n_exact": n_edges - n_approx,
            "n_fallback": n_fallback,
            "n_clamped": n_clamped,
            "approx_coverage": n_approx / n_edges if n_edges else 0.0,
            "fallback_ratio": (n_fallback / (n_approx + n_fallback)
                               if (n_approx + n_fallback) else 0.0),
            "clamp_ratio": n_clamped / n_approx if n_approx else 0.0,
        },
    }


# ---------------------------------------------------------------------------
# DOULION  (experiment/exp_gpu.py:396-450; src/doulion.py:11-84)
# ---------------------------------------------------------------------------

PCG64_MULT = 0x2360ED051FC65DA44385DF649FCCF645


def pcg64_doubles(state: int, inc: int, n: int) -> np.ndarray:
    """n doubles of numpy's PCG64 from (state, inc), restated with Python ints:
    s <- s * PCG64_MULT + inc (mod 2^128), x = rotr64(hi ^ lo, hi >> 58),
    u = (x >> 11) * 2^-53 (O'Neill's PCG XSL-RR 128/64; numpy's next_double)."""
    m128, m64 = (1 << 128) - 1, (1 << 64) - 1
    out = np.empty(n, dtype=np.float64)
    for i in range(n):
        state = (state * PCG64_MULT + inc) & m128
        hi, lo = state >> 64, state & m64
        x, r = hi ^ lo, hi >> 58
        x = ((x >> r) | (x << ((64 - r) & 63))) & m64
        out[i] = (x >> 11) * (1.0 / 9007199254740992.0)
    return out


def doulion_forward(row_ptr, col_idx, q: float, n_seeds: int = 10, base_seed: int = 42):
    """``_doulion_forward`` (experiment/exp_gpu.py:396-450) on a forward CSR.

    Edge e (CSR u-major order, the reference's edge_u/edge_v order) is kept
    iff ``np.random.default_rng(seed).random(E)[e] < q``; the sample keeps the
    orientation and is counted exactly; estimate = count / q^3.  Returns
    (estimates, sampled_counts).
    """
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    col_idx = np.asarray(col_idx)
    E = len(col_idx)
    eu = np.repeat(np.arange(len(row_ptr) - 1, dtype=np.int64), np.diff(row_ptr))
    estimates, counts = [], []
    for seed in range(base_seed, base_seed + n_seeds):
        mask = np.random.default_rng(seed).random(E) < q
        su, sv = eu[mask], col_idx[mask].astype(np.int64)
        rp = np.zeros(len(row_ptr), dtype=np.int64)
        np.cumsum(np.bincount(su, minlength=len(row_ptr) - 1), out=rp[1:])
        c = exact_count_csr(rp, sv) if len(sv) else 0
        counts.append(int(c))
        estimates.append(c / (q ** 3))
    return estimates, counts


def relative_error(estimate: float, truth: int) -> float:
    """``|est - truth| / truth`` with the zero conventions of src/utils.py:36-43."""
    if truth == 0:
        return float("inf") if estimate != 0 else 0.0
    return abs(estimate - truth) / truth


# ---------------------------------------------------------------------------
# Reference known-answer graph (cpu_exp/triangle_hll_experiment.py:1054-1069)
# ---------------------------------------------------------------------------

SELF_TEST_EDGES = np.array([[0, 1], [0, 2], [0, 3], [1, 2], [1, 3], [2, 3], [3, 4], [4, 5]],
                           dtype=np.int32)
SELF_TEST_TRIANGLES = 4
