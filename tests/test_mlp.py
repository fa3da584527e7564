"""Learned rank predictor: the FP64 numpy oracle is pinned by finite
differences (CPU); the GPU forward / backward / Adam kernels are checked
against it (GPU)."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def _setup(seed=0, n=300, n_in=6, hidden=16, layers=4):
    rng = np.random.default_rng(seed)
    d = M.dims(n_in, hidden, layers)
    npar = sum(d[l] * d[l + 1] + d[l + 1] for l in range(len(d) - 1))
    p = rng.standard_normal(npar) * 0.3
    X = rng.standard_normal((n, n_in)) * 2 + 1
    mu, sigma = X.mean(0), X.std(0)
    t = 20 + 10 * np.tanh(X[:, 0]) + 3 * X[:, 1] ** 2
    return d, p, X, t, mu, sigma


def test_oracle_gradient_matches_finite_differences():
    d, p, X, t, mu, sigma = _setup()
    loss, g = M.loss_grad(p, d, X, t, mu, sigma, 20.0, 10.0)
    rng = np.random.default_rng(1)
    idx = rng.choice(len(p), 40, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (M.loss_grad(p + e, d, X, t, mu, sigma, 20.0, 10.0)[0] -
              M.loss_grad(p - e, d, X, t, mu, sigma, 20.0, 10.0)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(g[i]))


@pytest.mark.gpu
def test_gpu_mlp_forward_backward_match_oracle(bfgpu):
    print("[case] mlp forward/backward vs oracle", flush=True)
    d, p, X, t, mu, sigma = _setup(n=1000, hidden=128)
    net = bfgpu.RankPredictor(6, 128, 4, seed=0)
    assert net.n_params() == len(p)
    net.set_params(p.astype(np.float32))
    net.set_norm(mu, sigma, 20.0, 10.0)
    p32 = net.get_params().astype(np.float64)
    y = net.forward(X.astype(np.float32))
    yo, _ = M.forward(p32, d, X.astype(np.float32), mu.astype(np.float32), sigma.astype(np.float32))
    np.testing.assert_allclose(y, yo * 10.0 + 20.0, rtol=1e-4, atol=1e-3)
    loss, g = net.loss_grad(X.astype(np.float32), t.astype(np.float32))
    lo, go = M.loss_grad(p32, d, X.astype(np.float32), t.astype(np.float32), mu.astype(np.float32),
                         sigma.astype(np.float32), 20.0, 10.0)
    assert abs(loss - lo) < 1e-4 * max(1.0, lo)
    assert np.linalg.norm(g - go) / np.linalg.norm(go) < 1e-4


@pytest.mark.gpu
def test_gpu_mlp_training_reduces_loss(bfgpu):
    print("[case] mlp training", flush=True)
    d, p, X, t, mu, sigma = _setup(n=4000, hidden=64)
    net = bfgpu.RankPredictor(6, 64, 4, seed=1)
    net.set_norm(mu, sigma, float(t.mean()), float(t.std()))
    l0, _ = net.loss_grad(X.astype(np.float32), t.astype(np.float32))
    net.train(X, t, epochs=60, batch=256, lr=2e-3, seed=0)
    l1, _ = net.loss_grad(X.astype(np.float32), t.astype(np.float32))
    assert l1 < 0.05 * l0
