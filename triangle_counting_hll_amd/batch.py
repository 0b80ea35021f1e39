"""Batched exact counting of many graphs in one launch.

No reference counterpart: the reference counts one SNAP graph per run
(experiment/exp_gpu.py:632-935).  Here thousands of graphs of uneven size are
packed into ONE forward CSR over disjoint id ranges, counted by one exact
pipeline that credits each triangle to the graph of its center vertex, and --
for repeated counts -- captured once into a HIP graph and replayed with one
launch (``BatchCountGraph``).

The many-graph workload: 4,096 graphs, seed 3; half Erdos-Renyi with n
uniform in [200, 50,000] and p = 8/n, half R-MAT with scale uniform in
[8, 16] and edge factor 16 (``many_graphs_spec``).
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib, gpu_graph, gpu_hybrid
from .graphgen import _rmat_from_key, draws, er_pairs_count, rmat_thresholds, stream_key

RMAT_ABCD = (0.57, 0.19, 0.19, 0.05)


@dataclass
class BatchSpec:
    """Parameters of a batch of R-MAT / ER graphs (one counter stream per graph)."""

    kind: np.ndarray      # uint8 [G]: 0 = Erdos-Renyi, 1 = R-MAT
    param: np.ndarray     # int32 [G]: n (ER) or scale (R-MAT)
    n_ids: np.ndarray     # int64 [G]: id range of each graph
    n_pairs: np.ndarray   # int64 [G]: raw pairs generated for each graph
    keys: np.ndarray      # uint64 [G]: counter-stream key of each graph
    seed: int
    abcd: tuple = RMAT_ABCD

    @property
    def n_graphs(self) -> int:
        return len(self.kind)

    @property
    def id_off(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.n_ids)]).astype(np.int64)

    @property
    def pair_off(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.n_pairs)]).astype(np.int64)

    def subset(self, idx) -> "BatchSpec":
        idx = np.asarray(idx)
        return BatchSpec(self.kind[idx], self.param[idx], self.n_ids[idx], self.n_pairs[idx],
                         self.keys[idx], self.seed, self.abcd)


def many_graphs_spec(n_graphs: int = 4096, seed: int = 3, er_n=(200, 50_000),
                     er_mean_degree: float = 8.0, rmat_scale=(8, 16), rmat_edge_factor: int = 16,
                     scale_down: int = 0) -> BatchSpec:
    """The many-graph workload: half ER (p = mean_degree / n), half R-MAT.

    ``scale_down`` divides ER sizes by 2**scale_down and lowers R-MAT scales by
    it (minimum 4), for quick tests.
    """
    rng = np.random.default_rng(seed)
    kind = rng.permutation(np.repeat(np.array([0, 1], dtype=np.uint8),
                                     [n_graphs - n_graphs // 2, n_graphs // 2]))
    n_er = rng.integers(er_n[0], er_n[1] + 1, n_graphs) >> scale_down
    n_er = np.maximum(n_er, 8)
    scale = np.maximum(rng.integers(rmat_scale[0], rmat_scale[1] + 1, n_graphs) - scale_down, 4)
    param = np.where(kind == 0, n_er, scale).astype(np.int32)
    n_ids = np.where(kind == 0, n_er, 1 << scale).astype(np.int64)
    n_pairs = np.array([er_pairs_count(int(n), er_mean_degree / n) if k == 0
                        else rmat_edge_factor << int(s)
                        for k, n, s in zip(kind, n_er, scale)], dtype=np.int64)
    keys = np.array([stream_key(seed, g) for g in range(n_graphs)], dtype=np.uint64)
    return BatchSpec(kind, param, n_ids, n_pairs, keys, seed)


def generate_pairs_numpy(spec: BatchSpec) -> np.ndarray:
    """Host twin of the device generator (same pairs, global ids); small batches only."""
    out = []
    a, b, c, _ = spec.abcd
    off = spec.id_off
    for g in range(spec.n_graphs):
        if spec.kind[g] == 0:
            p = _er_from_key(int(spec.keys[g]), int(spec.param[g]), int(spec.n_pairs[g]))
        else:
            p = _rmat_from_key(int(spec.keys[g]), int(spec.param[g]), int(spec.n_pairs[g]),
                               a, b, c)
        out.append(p.astype(np.int64) + off[g])
    if not out:
        return np.zeros((0, 2), dtype=np.int32)
    return np.concatenate(out).astype(np.int32)


def _er_from_key(key: int, n: int, m: int) -> np.ndarray:
    r = draws(key, 0, m)
    u = ((r & np.uint64(0xFFFFFFFF)) * np.uint64(n)) >> np.uint64(32)
    v = ((r >> np.uint64(32)) * np.uint64(n)) >> np.uint64(32)
    return np.stack([u, v], axis=1).astype(np.int32)


@_lib.on_stream
def generate_pairs_gpu(spec: BatchSpec, stream=None):
    """All raw pairs of the batch on the device (int32 [total, 2], global ids)."""
    t = _lib.require_device()
    if int(spec.id_off[-1]) > np.iinfo(np.int32).max:
        raise ValueError("the batch's id ranges do not fit int32")
    dev = t.device("cuda")
    kind = t.from_numpy(spec.kind.astype(np.uint8)).to(dev)
    param = t.from_numpy(spec.param.astype(np.int32)).to(dev)
    id_off = t.from_numpy(spec.id_off).to(dev)
    pair_off = t.from_numpy(spec.pair_off).to(dev)
    keys = t.from_numpy(spec.keys.view(np.int64)).to(dev)
    total = int(spec.pair_off[-1])
    pairs = t.empty((total, 2), dtype=t.int32, device=dev)
    t_a, t_ab, t_abc = rmat_thresholds(*spec.abcd[:3])
    _lib.check(_lib.load().tc_generate_pairs(
        _lib.ptr(kind), _lib.ptr(param), _lib.ptr(id_off), _lib.ptr(pair_off), _lib.ptr(keys),
        spec.n_graphs, t_a, t_ab, t_abc, _lib.ptr(pairs), total, _lib.stream_ptr(stream)))
    return pairs


def rmat_gpu(scale: int, edge_factor: int, seed: int, abcd=RMAT_ABCD):
    """Device twin of ``graphgen.rmat(scale, edge_factor, seed=seed)`` (stream 0)."""
    spec = BatchSpec(np.array([1], np.uint8), np.array([scale], np.int32),
                     np.array([1 << scale], np.int64),
                     np.array([int(edge_factor) << int(scale)], np.int64),
                     np.array([stream_key(seed, 0)], np.uint64), seed, tuple(abcd))
    return generate_pairs_gpu(spec)


@dataclass
class BatchedCSR:
    """One forward CSR holding every graph of a batch, plus the graph of each vertex."""

    csr: gpu_graph.DeviceCSR
    graph_of: object       # int32 [n] device
    n_graphs: int
    edge_off: np.ndarray   # int64 [G + 1]: graph g owns oriented edges [edge_off[g], edge_off[g+1])
    vert_off: np.ndarray   # int64 [G + 1]: graph g owns dense vertices [vert_off[g], vert_off[g+1])


@_lib.on_stream
def build_batch(spec: BatchSpec, pairs=None, stream=None) -> BatchedCSR:
    """Generate (unless ``pairs`` is given) and build the packed CSR on the device."""
    t = _lib.require_device()
    if pairs is None:
        pairs = generate_pairs_gpu(spec, stream)
    csr = gpu_graph.build_csr_gpu(pairs, row_ptr_dtype="int32", stream=stream)
    del pairs
    G = spec.n_graphs
    graph_of = t.empty(csr.n_nodes, dtype=t.int32, device="cuda")
    id_off = t.from_numpy(spec.id_off).cuda()
    _lib.check(_lib.load().tc_vertex_graph_ids(_lib.ptr(csr.node_ids), csr.n_nodes,
                                               _lib.ptr(id_off), G, _lib.ptr(graph_of),
                                               _lib.stream_ptr(stream)))
    # vertex and edge ranges of each graph (graphs are contiguous in dense ids)
    counts = t.bincount(graph_of.long(), minlength=G).cpu().numpy()
    vert_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    rp = csr.row_ptr.cpu().numpy().astype(np.int64)
    edge_off = rp[vert_off]
    return BatchedCSR(csr, graph_of, G, edge_off, vert_off)


@_lib.on_stream
def batch_exact_count(b: BatchedCSR, params: Optional[dict] = None, counts=None, stream=None):
    """Per-graph exact counts (device uint64 [G] viewed as int64), one eager pipeline."""
    t = _lib.require_device()
    if counts is None:
        counts = t.zeros(b.n_graphs, dtype=t.int64, device="cuda")
    p = gpu_hybrid._exact_params(params)
    rp, is64 = _lib.row_ptr_arg(b.csr.row_ptr)
    _lib.check(_lib.load().tc_batch_exact_count(
        _lib.ptr(rp), is64, _lib.ptr(b.csr.col_idx), b.csr.n_nodes, _lib.ptr(b.csr.edge_u),
        _lib.ptr(b.csr.edge_v), b.csr.n_oriented, _lib.ptr(b.graph_of), b.n_graphs,
        _lib.ptr(counts), ctypes.byref(p) if p is not None else None, _lib.stream_ptr(stream)))
    return counts


class BatchCountGraph:
    """The batched pipeline captured once into a HIP graph; ``launch()`` replays it.

    After a launch completes, ``counts`` holds the per-graph counts.
    """

    @_lib.on_stream
    def __init__(self, b: BatchedCSR, params: Optional[dict] = None, stream=None):
        t = _lib.require_device()
        self._b = b
        self.counts = t.zeros(b.n_graphs, dtype=t.int64, device="cuda")
        self._rp, is64 = _lib.row_ptr_arg(b.csr.row_ptr)
        p = gpu_hybrid._exact_params(params)
        h = ctypes.c_void_p()
        ws = ctypes.c_int64()
        _lib.check(_lib.load().tc_batch_graph_create(
            _lib.ptr(self._rp), is64, _lib.ptr(b.csr.col_idx), b.csr.n_nodes,
            _lib.ptr(b.csr.edge_u), _lib.ptr(b.csr.edge_v), b.csr.n_oriented,
            _lib.ptr(b.graph_of), b.n_graphs, _lib.ptr(self.counts),
            ctypes.byref(p) if p is not None else None, _lib.stream_ptr(stream),
            ctypes.byref(h), ctypes.byref(ws)))
        self._h = h
        self.workspace_bytes = ws.value

    @_lib.on_stream
    def launch(self, stream=None):
        _lib.check(_lib.load().tc_batch_graph_launch(self._h, _lib.stream_ptr(stream)))
        return self.counts

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().tc_batch_graph_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class GraphCSR:
    """One graph of a batch as its own forward CSR (dense ids from 0)."""

    row_ptr: object   # int32 [n_g + 1] device
    col_idx: object   # int32 [E_g] device
    edge_u: object
    edge_v: object
    v0: int           # first vertex of the graph in the packed CSR


def per_graph_csrs(b: BatchedCSR) -> list:
    """Every graph of the batch as the CSR a one-graph-per-run caller holds
    (the reference counts one graph per process, experiment/exp_gpu.py:632-935):
    its rows and ids rebased to start at 0.  Device slices, built once."""
    t = _lib.require_device()
    rp = b.csr.row_ptr.to(t.int64)
    out = []
    for g in range(b.n_graphs):
        v0, v1 = int(b.vert_off[g]), int(b.vert_off[g + 1])
        e0, e1 = int(b.edge_off[g]), int(b.edge_off[g + 1])
        out.append(GraphCSR((rp[v0:v1 + 1] - e0).to(t.int32), b.csr.col_idx[e0:e1] - v0,
                            b.csr.edge_u[e0:e1] - v0, b.csr.edge_v[e0:e1] - v0, v0))
    return out


def plain_per_graph_counts(b: BatchedCSR, params: Optional[dict] = None, graphs=None):
    """The plain form: one ``gpu_exact_triangle_count`` call per graph, each on
    that graph's own CSR (``per_graph_csrs``, built before timing).

    Returns ``(counts int64[G], seconds, device_ms)``: host wall time of the
    loop (each call synchronises, as the reference launcher does,
    src/gpu_hybrid.py:346-354) and the sum over graphs of the device time
    between HIP events recorded around each call on the stream.
    """
    t = _lib.require_device()
    graphs = graphs if graphs is not None else per_graph_csrs(b)
    out = np.zeros(b.n_graphs, dtype=np.int64)
    ev = [(t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True))
          for _ in range(b.n_graphs)]
    t.cuda.synchronize()
    t0 = time.perf_counter()
    for g, c in enumerate(graphs):
        if c.col_idx.numel():
            ev[g][0].record()
            out[g], _ = gpu_hybrid.gpu_exact_triangle_count(c.row_ptr, c.col_idx, c.edge_u,
                                                            c.edge_v, params=params)
            ev[g][1].record()
    secs = time.perf_counter() - t0
    t.cuda.synchronize()
    dev_ms = sum(a.elapsed_time(z) for (a, z), c in zip(ev, graphs) if c.col_idx.numel())
    return out, secs, dev_ms


def plain_per_graph_hybrid(b: BatchedCSR, gamma: int = 1024, p: int = 12, graphs=None):
    """The plain fixed-threshold hybrid, one graph at a time.

    One sketch build over the batch (vertices with forward degree >= gamma,
    compact register slots; the reference experiment passes theta = gamma,
    experiment/exp_gpu.py:556-557), then one ``gpu_hybrid_triangle_count`` call
    per graph on that graph's own CSR (src/gpu_hybrid.py:357-431: split exact
    and HLL kernels, synchronised).  Returns ``(estimates float64[G], seconds,
    build_seconds)``.
    """
    from . import gpu_hll

    t = _lib.require_device()
    graphs = graphs if graphs is not None else per_graph_csrs(b)
    deg = b.csr.degrees
    t.cuda.synchronize()
    t0 = time.perf_counter()
    sk = deg >= gamma
    n_sk = int(sk.sum().item())
    slot = t.full_like(deg, -1)
    slot[sk] = t.arange(n_sk, dtype=t.int32, device=deg.device)
    regs = t.zeros(max(n_sk, 1) << p, dtype=t.uint8, device=deg.device)
    has = t.zeros_like(deg, dtype=t.uint8)
    gpu_hll.build_sketches(b.csr.row_ptr, b.csr.col_idx, b.csr.n_nodes, gamma, p, regs, has,
                           slot=slot, accumulate=False)
    t.cuda.synchronize()
    build_s = time.perf_counter() - t0
    out = np.zeros(b.n_graphs, dtype=np.float64)
    t0 = time.perf_counter()
    for g, c in enumerate(graphs):
        if c.col_idx.numel():
            v0, v1 = int(b.vert_off[g]), int(b.vert_off[g + 1])
            out[g], _, _ = gpu_hybrid.gpu_hybrid_triangle_count(
                c.row_ptr, c.col_idx, c.edge_u, c.edge_v, deg[v0:v1], regs, has[v0:v1], gamma,
                p, slot=slot[v0:v1])
    return out, build_s + time.perf_counter() - t0, build_s
