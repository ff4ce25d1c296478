"""Float64 numpy restatement of the plan-selector MLP (test infrastructure).

Parameter layout as the engine's: for each layer W[in][out] (out fastest)
then b[out]; hidden layers ReLU, output linear; inputs normalized as
(x - mean) / std, loss = mean((f(x) - (y - ymean) / ystd)^2).
"""
import numpy as np


def layers(nin, hidden, width):
    dims = [nin] + [width] * hidden + [1]
    out, off = [], 0
    for a, c in zip(dims[:-1], dims[1:]):
        out.append((off, off + a * c, a, c))
        off += a * c + c
    return out, off


def unpack(p, nin, hidden, width):
    ls, _ = layers(nin, hidden, width)
    return [(p[w:b].reshape(a, c), p[b:b + c]) for w, b, a, c in ls]


def forward(p, x, nin, hidden, width):
    """Returns outputs and the per-layer (input, pre-activation) cache."""
    h = x
    cache = []
    Ws = unpack(p, nin, hidden, width)
    for li, (W, b) in enumerate(Ws):
        z = h @ W + b
        cache.append((h, z))
        h = z if li == len(Ws) - 1 else np.maximum(z, 0.0)
    return h[:, 0], cache


def loss_and_grad(p, x, y, nin, hidden, width):
    p = np.asarray(p, np.float64)
    out, cache = forward(p, x, nin, hidden, width)
    n = len(y)
    loss = np.mean((out - y) ** 2)
    dz = (2.0 * (out - y) / n)[:, None]
    ls, total = layers(nin, hidden, width)
    Ws = unpack(p, nin, hidden, width)
    g = np.zeros(total)
    for li in range(len(ls) - 1, -1, -1):
        w0, b0, a, c = ls[li]
        h, _ = cache[li]
        g[w0:b0] = (h.T @ dz).ravel()
        g[b0:b0 + c] = dz.sum(0)
        if li > 0:
            dh = dz @ Ws[li][0].T
            dz = dh * (cache[li - 1][1] > 0)
    return loss, g
