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


def test_oracle_matches_torch_linear_relu_autograd():
    """The oracle's layer order, ReLU placement, standardisation and loss
    against an independent torch model: nn.Linear layers loaded from the
    flat parameter vector (W_l row-major out x in, then b_l), ReLU between
    layers, MSE against (t - shift) / scale, gradients by autograd (FP64)."""
    torch = pytest.importorskip("torch")
    d, p, X, t, mu, sigma = _setup(seed=3, n=257, hidden=12, layers=4)
    shift, scale = 20.0, 10.0
    layers, o = [], 0
    for l in range(len(d) - 1):
        lin = torch.nn.Linear(d[l], d[l + 1]).double()
        with torch.no_grad():
            lin.weight.copy_(torch.from_numpy(p[o:o + d[l] * d[l + 1]].reshape(d[l + 1], d[l])))
            o += d[l] * d[l + 1]
            lin.bias.copy_(torch.from_numpy(p[o:o + d[l + 1]]))
            o += d[l + 1]
        layers.append(lin)
    assert o == len(p)
    mods = []
    for l, lin in enumerate(layers):
        mods.append(lin)
        if l < len(layers) - 1:
            mods.append(torch.nn.ReLU())
    net = torch.nn.Sequential(*mods)
    xt = (torch.from_numpy(X) - torch.from_numpy(mu)) / torch.from_numpy(sigma)
    yt = net(xt)[:, 0]
    target = (torch.from_numpy(t) - shift) / scale
    loss_t = torch.mean((yt - target) ** 2)
    loss_t.backward()
    g_t = torch.cat([torch.cat([lin.weight.grad.reshape(-1), lin.bias.grad]) for lin in layers]).numpy()
    y_o, _ = M.forward(p, d, X, mu, sigma)
    loss_o, g_o = M.loss_grad(p, d, X, t, mu, sigma, shift, scale)
    np.testing.assert_allclose(y_o, yt.detach().numpy(), rtol=1e-12, atol=1e-12)
    assert abs(loss_o - loss_t.item()) < 1e-12 * max(1.0, loss_o)
    np.testing.assert_allclose(g_o, g_t, rtol=1e-10, atol=1e-12)
    # a plausible slip fails: ReLU after the last layer as well
    bad = torch.nn.Sequential(*(mods + [torch.nn.ReLU()]))
    assert not np.allclose(y_o, bad(xt)[:, 0].detach().numpy(), rtol=1e-6, atol=1e-9)


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
