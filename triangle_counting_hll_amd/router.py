"""Learned exact-versus-sketch router for the hybrid count.

The reference routes an oriented edge to the HLL path iff
``min(deg+u, deg+v) >= gamma`` and both endpoints have sketches
(src/hybrid.py:107-112, src/gpu_hybrid.py:80-89).  Here a 6 -> 64 -> 1 MLP
(ReLU, sigmoid) makes that decision per edge instead:

* candidates: edges whose two endpoints have sketches (forward degree >=
  theta) -- the edges the degree rule with gamma = theta would send to HLL;
* features (``tc_router_features``): log2(1 + deg u), log2(1 + deg v), their
  difference, log2(1 + sketch cardinality) of u and of v, and log2(1 +
  union estimate) taken from a 2^-(p - s) register sample (2 * 2^s bytes per
  edge instead of 2 * 2^p);
* labels: 1 iff the full-precision HLL support of the edge lies within
  ``max(abs_tol, rel_tol * exact)`` of the exact support, both produced by
  this package's own kernels (``tc_edge_supports``, ``tc_union_estimate``);
* training: FP32, mean weighted cross-entropy, Adam, all on the GPU through
  ``tc_mlp_grad`` / ``tc_mlp_adam`` (gradients checked against the numpy
  oracle in tests);
* inference: BF16 (``tc_mlp_route_bf16``), threshold chosen on the training
  graph so that the HLL-routed edges keep the total relative error within a
  budget;
* optional cost gate: an edge may only take HLL when its exact work is larger
  than the sketch read, ``min(deg+u, deg+v) >= min_cost_degree``.

Edges the router does not send to HLL -- and every non-candidate -- take the
exact path, which stays the fallback and the judge of the router's error.
"""

from __future__ import annotations

import ctypes
import json
import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib, gpu_hll, gpu_hybrid

N_IN = 6
N_HIDDEN = 64
N_PARAMS = N_HIDDEN * N_IN + N_HIDDEN + N_HIDDEN + 1  # 513


def pack_params(W1, b1, w2, b2) -> np.ndarray:
    """Flat float32 parameter vector in the kernels' layout."""
    return np.concatenate([np.asarray(W1, np.float32).reshape(-1), np.asarray(b1, np.float32),
                           np.asarray(w2, np.float32),
                           np.asarray(b2, np.float32).reshape(-1)]).astype(np.float32)


def unpack_params(flat) -> dict:
    flat = np.asarray(flat, dtype=np.float32)
    o1 = N_HIDDEN * N_IN
    return {"W1": flat[:o1].reshape(N_HIDDEN, N_IN), "b1": flat[o1:o1 + N_HIDDEN],
            "w2": flat[o1 + N_HIDDEN:o1 + 2 * N_HIDDEN], "b2": flat[o1 + 2 * N_HIDDEN:]}


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform first layer, Glorot-uniform output layer, zero biases."""
    rng = np.random.default_rng(seed)
    l1 = math.sqrt(6.0 / N_IN)
    l2 = math.sqrt(6.0 / (N_HIDDEN + 1))
    return pack_params(rng.uniform(-l1, l1, (N_HIDDEN, N_IN)), np.zeros(N_HIDDEN),
                       rng.uniform(-l2, l2, N_HIDDEN), np.zeros(1))


@dataclass
class RouterModel:
    params: np.ndarray                 # float32 [513]
    norm: np.ndarray                   # float32 [12]: mean[6], inv_std[6]
    threshold: float = 0.5
    p: int = 12
    sample_bits: int = 8
    theta: int = 128
    min_cost_degree: int = 0
    meta: dict = field(default_factory=dict)

    def to_json(self) -> str:
        return json.dumps({"params": self.params.tolist(), "norm": self.norm.tolist(),
                           "threshold": self.threshold, "p": self.p,
                           "sample_bits": self.sample_bits, "theta": self.theta,
                           "min_cost_degree": self.min_cost_degree, "meta": self.meta})

    @staticmethod
    def from_json(text: str) -> "RouterModel":
        d = json.loads(text)
        return RouterModel(np.array(d["params"], np.float32), np.array(d["norm"], np.float32),
                           d["threshold"], d["p"], d["sample_bits"], d["theta"],
                           d["min_cost_degree"], d.get("meta", {}))


# ---------------------------------------------------------------------------
# Device building blocks
# ---------------------------------------------------------------------------

@dataclass
class Sketches:
    """Compact sketches of the vertices with forward degree >= theta."""

    p: int
    theta: int
    slot: object    # int32 [n], -1 without sketch
    regs: object    # uint8 [n_sketched * 2^p]
    has: object     # uint8 [n]
    card: object    # float32 [n]: single-sketch estimate (0 without sketch)
    n_sketched: int


def build_sketches(csr, theta: int, p: int) -> Sketches:
    t = _lib.require_device()
    deg = csr.degrees
    sk = deg >= theta
    n_sk = int(sk.sum().item())
    slot = t.full_like(deg, -1)
    slot[sk] = t.arange(n_sk, dtype=t.int32, device=deg.device)
    regs = t.zeros(max(n_sk, 1) << p, dtype=t.uint8, device=deg.device)
    has = t.zeros(deg.numel(), dtype=t.uint8, device=deg.device)
    gpu_hll.build_sketches(csr.row_ptr, csr.col_idx, csr.n_nodes, theta, p, regs, has, slot=slot,
                           accumulate=False)
    card = t.zeros(deg.numel(), dtype=t.float32, device=deg.device)
    if n_sk:
        ids = t.nonzero(sk).flatten().to(t.int32)
        est = gpu_hll.union_estimate_pairs(regs, ids, ids, p, slot=slot)
        card[ids.long()] = est.float()
    return Sketches(p, theta, slot, regs, has, card, n_sk)


def candidate_edges(csr, sk: Sketches, min_cost_degree: int = 0):
    """Edge ids whose endpoints both have sketches (and pass the cost gate)."""
    t = _lib.torch()
    du = csr.degrees[csr.edge_u.long()]
    dv = csr.degrees[csr.edge_v.long()]
    ok = (sk.has[csr.edge_u.long()] == 1) & (sk.has[csr.edge_v.long()] == 1)
    if min_cost_degree > 0:
        ok &= t.minimum(du, dv) >= min_cost_degree
    return t.nonzero(ok).flatten().to(t.int32)


@_lib.on_stream
def features(csr, sk: Sketches, cand, sample_bits: int = 8, stream=None):
    t = _lib.require_device()
    n = cand.numel()
    f = t.empty((n, N_IN), dtype=t.float32, device=cand.device)
    _lib.check(_lib.load().tc_router_features(
        _lib.ptr(csr.degrees), _lib.ptr(csr.edge_u), _lib.ptr(csr.edge_v), _lib.ptr(cand), n,
        _lib.ptr(sk.regs), _lib.ptr(sk.slot), sk.p, int(sample_bits), _lib.ptr(sk.card),
        _lib.ptr(f), _lib.stream_ptr(stream)))
    return f


def edge_errors(csr, sk: Sketches, cand):
    """(exact support, clamped full-precision HLL support) of each candidate edge."""
    t = _lib.torch()
    eu = csr.edge_u[cand.long()]
    ev = csr.edge_v[cand.long()]
    exact = gpu_hybrid.edge_supports(csr.row_ptr, csr.col_idx, eu, ev).double()
    union = gpu_hll.union_estimate_pairs(sk.regs, eu, ev, sk.p, slot=sk.slot)
    du = csr.degrees[eu.long()].double()
    dv = csr.degrees[ev.long()].double()
    s = t.clamp(du + dv - union, min=t.zeros_like(union), max=t.minimum(du, dv))
    return exact, s


def make_labels(exact, s_hll, rel_tol: float = 0.05, abs_tol: float = 1.0):
    """1.0 where the HLL support is within max(abs_tol, rel_tol * exact) of exact."""
    t = _lib.torch()
    tol = t.clamp(exact * rel_tol, min=abs_tol)
    return ((s_hll - exact).abs() <= tol).float()


def feature_norm(feats) -> np.ndarray:
    f = feats.double()
    mean = f.mean(0)
    std = f.std(0).clamp(min=1e-6)
    return np.concatenate([mean.cpu().numpy(), (1.0 / std).cpu().numpy()]).astype(np.float32)


# ---------------------------------------------------------------------------
# Training (GPU kernels) and inference
# ---------------------------------------------------------------------------

@_lib.on_stream
def mlp_forward(params, norm, feats, stream=None):
    t = _lib.require_device()
    pr = _lib.as_device(params, t.float32)
    nm = _lib.as_device(norm, t.float32)
    f = _lib.as_device(feats, t.float32)
    out = t.empty(f.shape[0], dtype=t.float32, device=f.device)
    _lib.check(_lib.load().tc_mlp_forward(_lib.ptr(pr), _lib.ptr(nm), _lib.ptr(f), f.shape[0],
                                          _lib.ptr(out), _lib.stream_ptr(stream)))
    return out


@_lib.on_stream
def mlp_grad(params, norm, feats, target, pos_weight: float = 1.0, stream=None):
    """(grad float32 [513], loss float32 [1]) of the mean weighted BCE."""
    t = _lib.require_device()
    pr = _lib.as_device(params, t.float32)
    nm = _lib.as_device(norm, t.float32)
    f = _lib.as_device(feats, t.float32)
    tg = _lib.as_device(target, t.float32)
    g = t.empty(N_PARAMS, dtype=t.float32, device=f.device)
    loss = t.empty(1, dtype=t.float32, device=f.device)
    _lib.check(_lib.load().tc_mlp_grad(_lib.ptr(pr), _lib.ptr(nm), _lib.ptr(f), _lib.ptr(tg),
                                       f.shape[0], float(pos_weight), _lib.ptr(g),
                                       _lib.ptr(loss), _lib.stream_ptr(stream)))
    return g, loss


@_lib.on_stream
def mlp_adam(params, grad, m, v, step: int, lr: float = 1e-3, beta1: float = 0.9,
             beta2: float = 0.999, eps: float = 1e-8, stream=None):
    _lib.check(_lib.load().tc_mlp_adam(_lib.ptr(params), _lib.ptr(grad), _lib.ptr(m), _lib.ptr(v),
                                       params.numel(), int(step), lr, beta1, beta2, eps,
                                       _lib.stream_ptr(stream)))


@_lib.on_stream
def mlp_route_bf16(params, norm, feats, threshold: float, cand=None, route=None, prob=True,
                   stream=None):
    """BF16 inference; returns probabilities (when ``prob``) and fills ``route``."""
    t = _lib.require_device()
    pr = _lib.as_device(params, t.float32)
    nm = _lib.as_device(norm, t.float32)
    f = _lib.as_device(feats, t.float32)
    out = t.empty(f.shape[0], dtype=t.float32, device=f.device) if prob else None
    _lib.check(_lib.load().tc_mlp_route_bf16(_lib.ptr(pr), _lib.ptr(nm), _lib.ptr(f), f.shape[0],
                                             float(threshold), _lib.ptr(cand), _lib.ptr(route),
                                             _lib.ptr(out), _lib.stream_ptr(stream)))
    return out


def train(feats, target, epochs: int = 12, batch: int = 8192, lr: float = 1e-2,
          pos_weight: Optional[float] = None, seed: int = 0, norm=None):
    """Train the network on the GPU; returns (params float32 [513] host, norm, loss history).

    Adam with the learning rate halved after every quarter of the epochs.
    """
    t = _lib.require_device()
    f = _lib.as_device(feats, t.float32)
    tg = _lib.as_device(target, t.float32)
    n = f.shape[0]
    if norm is None:
        norm = feature_norm(f)
    nm = t.from_numpy(np.asarray(norm, np.float32)).cuda()
    if pos_weight is None:
        pos = float(tg.sum().item())
        pos_weight = float(min(max((n - pos) / max(pos, 1.0), 1.0), 50.0))
    params = t.from_numpy(init_params(seed)).cuda()
    m = t.zeros_like(params)
    v = t.zeros_like(params)
    gen = t.Generator(device="cuda")
    gen.manual_seed(seed)
    history = []
    step = 0
    for ep in range(epochs):
        order = t.randperm(n, device="cuda", generator=gen)
        lr_ep = lr * 0.5 ** (4 * ep // max(epochs, 1))
        for s in range(0, n, batch):
            idx = order[s:s + batch]
            g, loss = mlp_grad(params, nm, f[idx], tg[idx], pos_weight)
            step += 1
            mlp_adam(params, g, m, v, step, lr_ep)
        _, loss = mlp_grad(params, nm, f, tg, pos_weight)
        history.append(float(loss.item()))
    return params.cpu().numpy(), np.asarray(norm, np.float32), history


def choose_threshold(prob, err, truth_total: float, budget: float,
                     rule: str = "signed") -> float:
    """Lowest threshold whose HLL-routed prefix (by descending probability)
    keeps the error of the routed edges within budget * truth_total.

    ``rule="signed"``: every partial |sum of signed errors| stays within the
    budget (errors of both signs cancel).  ``rule="absolute"``: the sum of the
    absolute per-edge errors stays within it, a bound on the total error that
    does not rely on cancellation, so it carries over to other graphs with a
    similar error distribution (the calibration graph is not the graph
    counted)."""
    t = _lib.torch()
    if rule not in ("signed", "absolute"):
        raise ValueError(f"rule must be 'signed' or 'absolute', got {rule!r}")
    order = t.argsort(prob, descending=True)
    e = err[order].double()
    c = t.cumsum(e.abs() if rule == "absolute" else e, 0).abs()
    bad = (c > budget * max(truth_total, 1.0)).nonzero()
    k = int(bad[0].item()) if bad.numel() else int(c.numel())
    if k == 0:
        return 1.01  # route nothing
    return float(prob[order[k - 1]].item())


@_lib.on_stream
def route(model: RouterModel, csr, sk: Sketches, stream=None):
    """Per-edge uint8 route (1 = HLL) from the network; candidates only."""
    t = _lib.require_device()
    cand = candidate_edges(csr, sk, model.min_cost_degree)
    r = t.zeros(csr.n_oriented, dtype=t.uint8, device="cuda")
    if cand.numel():
        f = features(csr, sk, cand, model.sample_bits, stream)
        mlp_route_bf16(model.params, model.norm, f, model.threshold, cand=cand, route=r,
                       prob=False, stream=stream)
    return r, int(cand.numel())


@_lib.on_stream
def routed_count(model: RouterModel, csr, sk: Optional[Sketches] = None, per_edge=None,
                 fused: bool = True, stream=None):
    """Hybrid count with the learned router; returns (estimate, stats dict).

    ``fused`` (default): the network is evaluated per candidate edge inside
    the exact pipeline's edge pass (``tc_router_hybrid_count``), where the
    reference evaluates its degree predicate (src/gpu_hybrid.py:80-89).
    ``fused=False`` runs the same decisions as separate steps (candidate
    list, ``tc_router_features``, ``tc_mlp_route_bf16``, then the hybrid count
    with the route array).
    ``per_edge`` (float64 device [E], optional) receives every edge's support.
    ``stats["use"]`` is the per-edge route taken (uint8 device, 1 = HLL).
    """
    t = _lib.require_device()
    if sk is None:
        # only vertices that can be an endpoint of a candidate need a sketch
        need = max(model.theta, model.min_cost_degree)
        if int(csr.degrees.max().item() if csr.n_nodes else 0) < need:
            count = gpu_hybrid.exact_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v,
                                           stream=stream)
            if per_edge is not None:
                per_edge.copy_(gpu_hybrid.edge_supports(csr.row_ptr, csr.col_idx, csr.edge_u,
                                                        csr.edge_v).double())
            return float(count.item()), {
                "n_approx": 0, "n_exact": csr.n_oriented, "n_clamped": 0, "n_candidates": 0,
                "n_sketched": 0, "use": t.zeros(csr.n_oriented, dtype=t.uint8, device="cuda")}
        sk = build_sketches(csr, need, model.p)
    if not fused:
        r, n_cand = route(model, csr, sk, stream)
        c, s, fb = gpu_hybrid.hybrid_count(csr.row_ptr, csr.col_idx, csr.edge_u, csr.edge_v,
                                           csr.degrees, sk.regs, sk.has, 0, model.p, route=r,
                                           slot=sk.slot, per_edge=per_edge, stream=stream)
        st = s.cpu().numpy()
        return float(c.item()), {"n_approx": int(st[0]), "n_exact": int(st[1]),
                                 "n_clamped": int(st[2]), "n_candidates": n_cand,
                                 "n_sketched": sk.n_sketched, "use": r}
    c, s, n_cand, use = fused_count(model, csr, sk, per_edge=per_edge, stream=stream)
    st = s.cpu().numpy()
    return float(c.item()), {"n_approx": int(st[0]), "n_exact": int(st[1]),
                             "n_clamped": int(st[2]), "n_candidates": int(n_cand.item()),
                             "n_sketched": sk.n_sketched, "use": use}


@_lib.on_stream
def fused_count(model: RouterModel, csr, sk: Sketches, counts=None, stats=None, n_cand=None,
                use=None, per_edge=None, stream=None):
    """Enqueue ``tc_router_hybrid_count``: the network routes each candidate
    inside the exact pipeline's edge pass.  Returns the device accumulators
    ``(counts float64 [1], stats int64 [3], n_candidates int64 [1], use uint8 [E])``."""
    t = _lib.require_device()
    rp, is64 = _lib.row_ptr_arg(csr.row_ptr)
    dev = rp.device
    params = _lib.as_device(model.params, t.float32)
    norm = _lib.as_device(model.norm, t.float32)
    card = _lib.as_device(sk.card, t.float32)
    if params.numel() != N_PARAMS or norm.numel() != 2 * N_IN:
        raise ValueError("router parameters must be 513 floats and norm 12 floats")
    if counts is None:
        counts = t.zeros(1, dtype=t.float64, device=dev)
    if stats is None:
        stats = t.zeros(3, dtype=t.int64, device=dev)
    if n_cand is None:
        n_cand = t.zeros(1, dtype=t.int64, device=dev)
    if use is None:
        use = t.empty(max(csr.n_oriented, 1), dtype=t.uint8, device=dev)
    net = _lib.RouterNet(_lib.ptr(params), _lib.ptr(norm), _lib.ptr(card),
                         float(model.threshold), int(model.sample_bits),
                         int(model.min_cost_degree))
    _lib.check(_lib.load().tc_router_hybrid_count(
        _lib.ptr(rp), is64, _lib.ptr(csr.col_idx), csr.n_nodes, _lib.ptr(csr.edge_u),
        _lib.ptr(csr.edge_v), csr.n_oriented, _lib.ptr(csr.degrees), _lib.ptr(sk.regs),
        _lib.ptr(sk.slot), _lib.ptr(sk.has), int(sk.p), ctypes.byref(net), _lib.ptr(counts),
        _lib.ptr(stats), _lib.ptr(n_cand), _lib.ptr(use), _lib.ptr(per_edge),
        _lib.stream_ptr(stream)))
    return counts, stats, n_cand, use[: csr.n_oriented]
