"""Pins the learned-proposal oracle (oracle/mlp.py): analytic gradients
against central finite differences, covariance features against numpy.cov,
bfloat16 rounding against torch, Adam against torch.optim.Adam."""
import numpy as np
import torch

from oracle import mlp
from tests.hc_cases import blocks_from_texture


def small_batch(n=24, seed=0):
    blocks = blocks_from_texture(32, seed)[:n]
    x = mlp.features(blocks)
    rng = np.random.default_rng(seed)
    t = rng.random((n, 6)).astype(np.float32)
    return x, t


def test_gradients_match_finite_differences():
    x, t = small_batch()
    theta = mlp.init_params(3).astype(np.float64) * 0.5
    theta += np.random.default_rng(1).standard_normal(theta.shape) * 0.01
    y, _ = mlp.forward(theta, x, bf16_round=False)
    lab = mlp.conf_labels(y.astype(np.float32), t)
    _, g = mlp.loss_and_grads(theta, x, t, bf16_round=False, labels=lab)
    rng = np.random.default_rng(2)
    idx = np.concatenate([rng.choice(mlp.N_PARAMS, 60, replace=False),
                          np.arange(mlp.N_PARAMS - 7, mlp.N_PARAMS)])      # b3 (incl. the confidence bias)
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        lp, _ = mlp.loss_and_grads(tp, x, t, bf16_round=False, labels=lab)
        lm, _ = mlp.loss_and_grads(tm, x, t, bf16_round=False, labels=lab)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_covariance_features():
    blocks = blocks_from_texture(32, 5)[:50]
    f = mlp.features(blocks)
    for b in range(50):
        px = blocks[b, :, :3].astype(np.float64) / 255.0
        assert np.allclose(f[b, :48], (px - 0.5).ravel(), atol=1e-6)
        assert np.allclose(f[b, 48:].reshape(3, 3), np.cov(px.T, bias=True) * 4, atol=1e-6)


def test_bf16_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 10000),
                        np.array([1.00390625, 1.01171875, -2.00390625], np.float32)])   # ties
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(mlp.bf16(x), ref)


def test_adam_matches_torch():
    rng = np.random.default_rng(4)
    w = rng.standard_normal(50).astype(np.float64)
    m, v = np.zeros_like(w), np.zeros_like(w)
    tw = torch.tensor(w.copy(), requires_grad=True)
    opt = torch.optim.Adam([tw], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for step in range(1, 6):
        g = rng.standard_normal(50)
        mlp.adam(w, g, m, v, step)
        tw.grad = torch.tensor(g)
        opt.step()
    assert np.allclose(w, tw.detach().numpy(), rtol=1e-6, atol=1e-9)


def test_targets_brighter_first():
    rng = np.random.default_rng(5)
    lo = rng.integers(0, 65536, 500).astype(np.uint16)
    hi = rng.integers(0, 65536, 500).astype(np.uint16)
    t = mlp.targets_from_endpoints(lo, hi)
    sw = mlp.targets_from_endpoints(hi, lo)
    assert np.array_equal(t, sw)                       # order of the pair does not matter
    lum = lambda c: 2 * c[:, 0] * 31 * 33 / 4 + 4 * c[:, 1] * 63 * 65 / 16 + c[:, 2] * 31 * 33 / 4
    assert (lum(t[:, :3]) >= lum(t[:, 3:]) - 1e-3).all()
