"""Router kernels on the GPU against the numpy oracle, and what training buys."""

import math

import numpy as np
import pytest

from oracle import mlp_oracle as M
from oracle import tc_oracle as O
from triangle_counting_hll_amd import gpu_graph, router
from triangle_counting_hll_amd.graphgen import rmat

pytestmark = pytest.mark.gpu


def _problem(n=3000, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 6)).astype(np.float32) * 2 + 1
    t = (rng.random(n) < 0.35).astype(np.float32)
    norm = np.concatenate([X.mean(0), 1 / X.std(0)]).astype(np.float32)
    Xn = (X - norm[:6]) * norm[6:]
    return X, t, norm, Xn


@pytest.mark.parametrize("pos_weight", [1.0, 2.5])
def test_gradients_match_oracle(device, pos_weight):
    X, t, norm, Xn = _problem()
    flat = router.init_params(3)
    flat[384:448] = np.random.default_rng(1).normal(size=64) * 0.1
    g, loss = router.mlp_grad(flat, norm, X, t, pos_weight)
    p = router.unpack_params(flat)
    want = M.backward(p, Xn.astype(np.float64), t, pos_weight)
    want_flat = router.pack_params(want["W1"], want["b1"], want["w2"], want["b2"])
    np.testing.assert_allclose(g.cpu().numpy(), want_flat, rtol=2e-4, atol=2e-7)
    assert math.isclose(float(loss.item()), M.loss(p, Xn.astype(np.float64), t, pos_weight),
                        rel_tol=1e-5)
    # deterministic: same inputs, same bits
    g2, _ = router.mlp_grad(flat, norm, X, t, pos_weight)
    assert np.array_equal(g.cpu().numpy(), g2.cpu().numpy())


def test_forward_and_adam_match_oracle(device):
    import torch

    X, t, norm, Xn = _problem(777, 2)
    flat = router.init_params(5)
    p = router.unpack_params(flat)
    y = router.mlp_forward(flat, norm, X).cpu().numpy()
    np.testing.assert_allclose(y, M.forward(p, Xn.astype(np.float64))[3], rtol=1e-5, atol=1e-6)
    params = torch.from_numpy(flat.copy()).cuda()
    m = torch.zeros_like(params)
    v = torch.zeros_like(params)
    state = {}
    pd = {k: v_.astype(np.float64) for k, v_ in p.items()}
    for step in range(1, 4):
        g, _ = router.mlp_grad(params, norm, X, t)
        router.mlp_adam(params, g, m, v, step, lr=0.01)
        gd = router.unpack_params(g.cpu().numpy())
        pd = M.adam_step(pd, gd, state, step, lr=0.01)
        want = router.pack_params(pd["W1"], pd["b1"], pd["w2"], pd["b2"])
        np.testing.assert_allclose(params.cpu().numpy(), want, rtol=1e-5, atol=1e-6)
        pd = router.unpack_params(params.cpu().numpy())
        pd = {k: v_.astype(np.float64) for k, v_ in pd.items()}


def test_bf16_inference_matches_oracle(device):
    import torch

    X, t, norm, Xn = _problem(2000, 3)
    flat = router.init_params(9)
    prob = router.mlp_route_bf16(flat, norm, X, 0.5).cpu().numpy()
    want = M.forward_bf16(router.unpack_params(flat), Xn)
    np.testing.assert_allclose(prob, want, rtol=0, atol=2e-5)
    cand = torch.arange(1999, -1, -1, dtype=torch.int32, device="cuda")
    r = torch.zeros(2000, dtype=torch.uint8, device="cuda")
    router.mlp_route_bf16(flat, norm, X, 0.5, cand=cand, route=r, prob=False)
    got = r.cpu().numpy()[::-1].astype(bool)
    clear = np.abs(want - 0.5) > 1e-4
    assert np.array_equal(got[clear], (want >= 0.5)[clear])


def _graph(scale, seed, p, theta):
    pairs = rmat(scale, 16, seed=seed)
    csr = gpu_graph.build_csr_gpu(pairs)
    sk = router.build_sketches(csr, theta, p)
    return pairs, csr, sk


def test_features_match_oracle(device):
    p, s, theta = 10, 8, 24
    pairs, csr, sk = _graph(11, 2, p, theta)
    rp, col, fdeg, nid, eu, ev, _ = O.csr_from_pairs(pairs)
    cand = router.candidate_edges(csr, sk)
    f = router.features(csr, sk, cand, s).cpu().numpy()
    c = cand.cpu().numpy()
    deg = np.diff(rp)
    regs, has = O.build_registers(rp, col, p, theta)
    card = O.estimate_rows(regs)
    u, v = eu[c], ev[c]
    np.testing.assert_allclose(f[:, 0], np.log2(1 + deg[u]), rtol=1e-6)
    np.testing.assert_allclose(f[:, 2], np.log2(1 + deg[u]) - np.log2(1 + deg[v]), atol=1e-5)
    np.testing.assert_allclose(f[:, 3], np.log2(1 + card[u]), rtol=1e-5)
    sampled = O.union_estimate_rows(regs[u, : 1 << s], regs[v, : 1 << s]) * 2 ** (p - s)
    np.testing.assert_allclose(f[:, 5], np.log2(1 + sampled), rtol=1e-5)


def test_training_learns_and_beats_degree_rule(device):
    """What the network gains over the degree threshold: at the same number of
    HLL-routed edges it picks edges whose HLL supports are closer to exact."""
    import torch

    p, theta = 10, 16
    _, csr_a, sk_a = _graph(13, 5, p, theta)
    cand_a = router.candidate_edges(csr_a, sk_a)
    fa = router.features(csr_a, sk_a, cand_a)
    ex_a, hl_a = router.edge_errors(csr_a, sk_a, cand_a)
    ya = router.make_labels(ex_a, hl_a, rel_tol=0.05, abs_tol=1.0)
    params, norm, hist = router.train(fa, ya, epochs=8, batch=8192, lr=3e-3)
    assert hist[-1] < hist[0]

    # held-out graph
    _, csr_b, sk_b = _graph(13, 4, p, theta)
    cand_b = router.candidate_edges(csr_b, sk_b)
    fb = router.features(csr_b, sk_b, cand_b)
    ex_b, hl_b = router.edge_errors(csr_b, sk_b, cand_b)
    err = (hl_b - ex_b).abs().cpu().numpy()
    prob = router.mlp_route_bf16(params, norm, fb, 0.5).cpu().numpy()
    k = len(err) // 4
    picked_router = np.argsort(-prob, kind="stable")[:k]
    eu, ev = csr_b.edge_u[cand_b.long()], csr_b.edge_v[cand_b.long()]
    mind = torch.minimum(csr_b.degrees[eu.long()], csr_b.degrees[ev.long()]).cpu().numpy()
    picked_degree = np.argsort(-mind, kind="stable")[:k]
    assert err[picked_router].mean() < 0.8 * err[picked_degree].mean()


def test_routed_count_is_exact_when_nothing_routes(device):
    p, theta = 10, 8
    pairs, csr, sk = _graph(11, 3, p, theta)
    rp, col, *_ = O.csr_from_pairs(pairs)
    model = router.RouterModel(router.init_params(0), np.array([0] * 6 + [1] * 6, np.float32),
                               threshold=2.0, p=p, theta=theta)
    est, st = router.routed_count(model, csr, sk)
    assert st["n_approx"] == 0 and est == float(O.exact_count_csr(rp, col))
    model.threshold = -1.0  # every candidate takes HLL
    est2, st2 = router.routed_count(model, csr, sk)
    assert st2["n_approx"] == st2["n_candidates"] > 0
    back = router.RouterModel.from_json(model.to_json())
    assert np.array_equal(back.params, model.params) and back.threshold == -1.0


@pytest.mark.parametrize("threshold,gate,sample_bits", [(0.5, 0, 8), (0.3, 40, 8), (0.6, 0, 6),
                                                        (0.5, 0, 10), (0.5, 10_000, 8)])
def test_fused_routing_equals_separate_steps(device, threshold, gate, sample_bits):
    # the network evaluated inside the exact pass takes the same decisions as
    # features + BF16 inference as separate kernels, so the estimate, stats and
    # per-edge supports are identical
    import torch

    p, theta = 10, 16
    pairs, csr, sk = _graph(13, 6, p, theta)
    cand = router.candidate_edges(csr, sk)
    norm = router.feature_norm(router.features(csr, sk, cand, sample_bits))
    model = router.RouterModel(router.init_params(5), norm, threshold, p, sample_bits, theta,
                               min_cost_degree=gate)
    pa = torch.zeros(csr.n_oriented, dtype=torch.float64, device="cuda")
    pb = torch.zeros_like(pa)
    ea, sa = router.routed_count(model, csr, sk, per_edge=pa, fused=True)
    eb, sb = router.routed_count(model, csr, sk, per_edge=pb, fused=False)
    assert torch.equal(sa["use"], sb["use"])
    for k in ("n_approx", "n_exact", "n_clamped", "n_candidates"):
        assert sa[k] == sb[k], k
    assert ea == eb and torch.equal(pa, pb)
    if gate == 10_000:
        assert sa["n_candidates"] == 0 and sa["n_approx"] == 0
    else:
        assert 0 < sa["n_approx"] < sa["n_candidates"]
    # exact-routed supports are the oracle's
    rp, col, fdeg, nid, eu, ev, _ = O.csr_from_pairs(pairs)
    use = sa["use"].cpu().numpy().astype(bool)
    sup = O.edge_supports(rp, col, eu, ev)
    assert np.array_equal(pa.cpu().numpy()[~use], sup[~use].astype(np.float64))


def test_fused_router_rejects_bad_arguments(device):
    import torch

    p, theta = 10, 16
    _, csr, sk = _graph(10, 1, p, theta)
    model = router.RouterModel(router.init_params(0), np.array([0] * 6 + [1] * 6, np.float32),
                               0.5, p, 12, theta)
    with pytest.raises(ValueError, match="sample_bits"):
        router.fused_count(model, csr, sk)
    model.sample_bits = 8
    model.min_cost_degree = -1
    with pytest.raises(ValueError, match="min_cost_degree"):
        router.fused_count(model, csr, sk)
    model.min_cost_degree = 0
    model.params = model.params[:100]
    with pytest.raises(ValueError, match="513"):
        router.fused_count(model, csr, sk)
    torch.cuda.synchronize()
