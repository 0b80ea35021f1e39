"""Pins the router MLP oracle (oracle/mlp_oracle.py) before it judges the kernels."""

import numpy as np
import pytest

from oracle import mlp_oracle as M
from triangle_counting_hll_amd import router


def _data(n=64, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, M.N_IN))
    t = (rng.random(n) < 0.3).astype(np.float64)
    return X, t


@pytest.mark.parametrize("pos_weight", [1.0, 3.5])
def test_backward_matches_finite_differences(pos_weight):
    p = M.init_params(1)
    p = {k: v.astype(np.float64) for k, v in p.items()}
    p["b1"] = np.random.default_rng(2).normal(size=M.N_HIDDEN) * 0.1
    X, t = _data()
    g = M.backward(p, X, t, pos_weight)
    h = 1e-6
    rng = np.random.default_rng(3)
    for k in p:
        flat = p[k].reshape(-1)
        for idx in rng.choice(flat.size, min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp = M.loss(p, X, t, pos_weight)
            flat[idx] = old - h
            lm = M.loss(p, X, t, pos_weight)
            flat[idx] = old
            num = (lp - lm) / (2 * h)
            assert abs(num - g[k].reshape(-1)[idx]) < 1e-6 + 1e-5 * abs(num), (k, idx)


def test_forward_closed_form():
    p = {"W1": np.zeros((M.N_HIDDEN, M.N_IN), np.float32), "b1": np.ones(M.N_HIDDEN, np.float32),
         "w2": np.full(M.N_HIDDEN, 0.5, np.float32), "b2": np.array([-30.0], np.float32)}
    _, h, z, y = M.forward(p, np.zeros((3, M.N_IN)))
    assert np.allclose(h, 1.0) and np.allclose(z, 2.0)
    assert np.allclose(y, 1 / (1 + np.exp(-2.0)))


def test_adam_first_step_is_signed_lr():
    p = M.init_params(0)
    g = {k: np.full_like(v, 0.25, dtype=np.float64) for k, v in p.items()}
    g["b2"] = np.array([-2.0])
    out = M.adam_step(p, g, {}, 1, lr=0.01)
    # step 1: m_hat = g, v_hat = g^2 -> update = lr * g / (|g| + eps)
    assert np.allclose(out["W1"] - p["W1"], -0.01, atol=1e-9)
    assert np.allclose(out["b2"] - p["b2"], 0.01, atol=1e-9)


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-9, 1.0 + 2**-7 + 2**-8, -2.5, 0.0],
                 dtype=np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2**-7, 1.0 + 2**-6, -2.5, 0.0], dtype=np.float32)
    assert np.array_equal(M.to_bf16(x), want)


def test_bf16_forward_close_to_fp32():
    p = M.init_params(4)
    X, _ = _data(200, 5)
    y32 = M.forward(p, X)[3]
    y16 = M.forward_bf16(p, X)
    assert np.max(np.abs(y32 - y16)) < 0.05


def test_package_layout_round_trips():
    p = M.init_params(7)
    flat = router.pack_params(p["W1"], p["b1"], p["w2"], p["b2"])
    assert flat.shape == (router.N_PARAMS,) == (513,)
    back = router.unpack_params(flat)
    for k in p:
        assert np.array_equal(back[k].reshape(p[k].shape), p[k])
    # the package's initialiser matches the oracle's distribution contract
    q = router.unpack_params(router.init_params(7))
    assert np.abs(q["W1"]).max() <= np.sqrt(6.0 / 6) and not q["b1"].any()
