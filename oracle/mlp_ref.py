"""Proposer network restated in numpy -- TEST INFRASTRUCTURE ONLY.

The checker of the device kernels of fmm2d_amd/csrc/mlp.hip (forward pass,
weighted squared-error loss, its gradient, Adam).  Only tests/ may import it.
The network has no reference counterpart (the reference uses fixed nterms /
ndiv rules); this restatement is pinned by finite differences of the loss
(tests/test_oracle.py) and by closed forms on tiny cases.

Layout of the flat parameter vector: W1[64][nin], b1[64], W2[64][64], b2[64],
W3[nout][64], b3[nout].
"""
import numpy as np

H = 64


def layout(nin, nout):
    sizes = [H * nin, H, H * H, H, nout * H, nout]
    offs = np.cumsum([0] + sizes)
    return offs, int(offs[-1])


def unpack(p, nin, nout):
    offs, _ = layout(nin, nout)
    W1 = p[offs[0]:offs[1]].reshape(H, nin)
    b1 = p[offs[1]:offs[2]]
    W2 = p[offs[2]:offs[3]].reshape(H, H)
    b2 = p[offs[3]:offs[4]]
    W3 = p[offs[4]:offs[5]].reshape(nout, H)
    b3 = p[offs[5]:offs[6]]
    return W1, b1, W2, b2, W3, b3


def forward(p, x, nin, nout):
    W1, b1, W2, b2, W3, b3 = unpack(p, nin, nout)
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3.T + b3, (h1, h2)


def loss_grad(p, x, t, w, nin, nout):
    """loss = sum w (y - t)^2 / batch and d loss / d p (same layout as p)."""
    W1, b1, W2, b2, W3, b3 = unpack(p, nin, nout)
    batch = x.shape[0]
    y, (h1, h2) = forward(p, x, nin, nout)
    w = np.ones_like(y) if w is None else w
    r = y - t
    loss = float((w * r * r).sum() / batch)
    dy = 2.0 * w * r / batch
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    d2 = (dy @ W3) * (1.0 - h2 * h2)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (1.0 - h1 * h1)
    gW1 = d1.T @ x
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def adam(p, g, m, v, lr, b1, b2, eps, step):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def init(nin, nout, seed=0):
    """Glorot-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((nin, H), (H, H), (H, nout)):
        a = np.sqrt(6.0 / (fan_in + fan_out))
        parts.append(rng.uniform(-a, a, fan_out * fan_in))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts)
