"""CPU oracle of the learned router's network (test infrastructure only).

Restates, in numpy float64, the 6 -> 64 -> 1 MLP (ReLU, sigmoid) that the HIP
kernels train in FP32 and run in BF16: forward pass, weighted binary
cross-entropy, backward pass, one Adam step, and the BF16 inference rounding.
No reference counterpart exists (the reference has no learned router); these
are the textbook definitions (Kingma & Ba 2015 for Adam), pinned in
tests/test_router_oracle.py by finite differences and closed forms.
"""

from __future__ import annotations

import numpy as np

N_IN = 6
N_HIDDEN = 64


def init_params(seed: int = 0, n_in: int = N_IN, n_hidden: int = N_HIDDEN) -> dict:
    """He-uniform first layer, Glorot-uniform output layer, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    lim1 = np.sqrt(6.0 / n_in)
    lim2 = np.sqrt(6.0 / (n_hidden + 1))
    return {
        "W1": rng.uniform(-lim1, lim1, (n_hidden, n_in)).astype(np.float32),
        "b1": np.zeros(n_hidden, dtype=np.float32),
        "w2": rng.uniform(-lim2, lim2, n_hidden).astype(np.float32),
        "b2": np.zeros(1, dtype=np.float32),
    }


def forward(p: dict, X: np.ndarray):
    """Returns (pre-activation, hidden, logit, probability) in float64."""
    X = np.asarray(X, dtype=np.float64)
    a = X @ p["W1"].astype(np.float64).T + p["b1"].astype(np.float64)
    h = np.maximum(a, 0.0)
    z = h @ p["w2"].astype(np.float64) + float(p["b2"][0])
    y = 1.0 / (1.0 + np.exp(-z))
    return a, h, z, y


def softplus(x):
    return np.logaddexp(0.0, x)


def loss(p: dict, X, t, pos_weight: float = 1.0) -> float:
    """Mean weighted BCE: -(w t log y + (1 - t) log(1 - y)), computed from logits."""
    _, _, z, _ = forward(p, X)
    t = np.asarray(t, dtype=np.float64)
    return float(np.mean(pos_weight * t * softplus(-z) + (1.0 - t) * softplus(z)))


def backward(p: dict, X, t, pos_weight: float = 1.0) -> dict:
    """Gradients of ``loss`` (mean over the batch) in float64."""
    X = np.asarray(X, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    a, h, z, y = forward(p, X)
    B = X.shape[0]
    g = (pos_weight * t * (y - 1.0) + (1.0 - t) * y) / B      # dL/dz
    dw2 = h.T @ g
    db2 = np.array([g.sum()])
    dh = np.outer(g, p["w2"].astype(np.float64)) * (a > 0)
    dW1 = dh.T @ X
    db1 = dh.sum(axis=0)
    return {"W1": dW1, "b1": db1, "w2": dw2, "b2": db2}


def adam_step(p: dict, grads: dict, state: dict, step: int, lr: float = 1e-3,
              beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> dict:
    """One Adam update (bias-corrected, step counted from 1); returns new params."""
    out = {}
    for k in p:
        m = state.setdefault("m_" + k, np.zeros_like(p[k], dtype=np.float64))
        v = state.setdefault("v_" + k, np.zeros_like(p[k], dtype=np.float64))
        g = np.asarray(grads[k], dtype=np.float64)
        m = beta1 * m + (1 - beta1) * g
        v = beta2 * v + (1 - beta2) * g * g
        state["m_" + k], state["v_" + k] = m, v
        mh = m / (1 - beta1 ** step)
        vh = v / (1 - beta2 ** step)
        out[k] = p[k].astype(np.float64) - lr * mh / (np.sqrt(vh) + eps)
    return out


def to_bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even); returned as float32."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    rounding = ((b >> np.uint64(16)) & np.uint64(1)) + np.uint64(0x7FFF)
    r = ((b + rounding) >> np.uint64(16)) << np.uint64(16)
    return r.astype(np.uint32).view(np.float32)


def forward_bf16(p: dict, X) -> np.ndarray:
    """BF16 inference: weights, inputs and hidden activations rounded to bf16,
    products accumulated in float32, biases in float32."""
    Xb = to_bf16(np.asarray(X, dtype=np.float32)).astype(np.float32)
    W1 = to_bf16(p["W1"])
    w2 = to_bf16(p["w2"])
    a = (Xb @ W1.T).astype(np.float32) + p["b1"].astype(np.float32)
    h = to_bf16(np.maximum(a, 0.0))
    z = (h @ w2).astype(np.float32) + np.float32(p["b2"][0])
    return 1.0 / (1.0 + np.exp(-z.astype(np.float64)))
