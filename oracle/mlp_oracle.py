"""FP64 numpy oracle of the learned rank predictor (test infrastructure only).

Same parametrisation as the library's MLP: one flat parameter vector
W1, b1, ..., WL, bL with W_l row-major (out x in); inputs standardised as
(x - mu) / sigma; ReLU after every layer but the last; loss = mean squared
error between the output and (rank - shift) / scale.  Used by the tests to
check the GPU forward/backward kernels; its own gradient is pinned against
central finite differences.
"""
import numpy as np


def dims(n_in, hidden, n_layers):
    return [n_in] + [hidden] * (n_layers - 1) + [1]


def unpack(p, d):
    out, o = [], 0
    for l in range(len(d) - 1):
        W = p[o:o + d[l] * d[l + 1]].reshape(d[l + 1], d[l])
        o += d[l] * d[l + 1]
        b = p[o:o + d[l + 1]]
        o += d[l + 1]
        out.append((W, b))
    return out


def forward(p, d, X, mu, sigma):
    a = (np.asarray(X, np.float64) - mu) / sigma
    acts = [a]
    layers = unpack(np.asarray(p, np.float64), d)
    for l, (W, b) in enumerate(layers):
        z = a @ W.T + b
        a = z if l == len(layers) - 1 else np.maximum(z, 0.0)
        acts.append(a)
    return a[:, 0], acts


def loss_grad(p, d, X, t, mu, sigma, shift, scale):
    y, acts = forward(p, d, X, mu, sigma)
    n = len(y)
    r = y - (np.asarray(t, np.float64) - shift) / scale
    loss = float(np.mean(r * r))
    layers = unpack(np.asarray(p, np.float64), d)
    grads = []
    dz = (2.0 * r / n)[:, None]
    for l in range(len(layers) - 1, -1, -1):
        W, b = layers[l]
        grads.append((dz.T @ acts[l], dz.sum(axis=0)))
        if l > 0:
            dz = (dz @ W) * (acts[l] > 0)
    g = []
    for dW, db in reversed(grads):
        g.append(dW.reshape(-1))
        g.append(db)
    return loss, np.concatenate(g)
