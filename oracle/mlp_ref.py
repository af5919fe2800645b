"""CPU restatement of the warm-start network -- test infrastructure only.

Forward pass, weighted squared loss, backward pass and one Adam step in float64
numpy, with the parameter layout of the device network (qh_mlp.hip): W_0
(dims[1] x dims[0], row-major), b_0, W_1, b_1, ...; hidden layers use tanh,
the last layer is linear.  Loss = mean over the batch of sum_j 0.5 w_j (y_j - t_j)^2.

QuaC has no learned component, so there is no reference file to follow.  The
gradient is pinned by central finite differences (tests/test_oracle_mlp.py), and
the Adam step by its published closed-form properties (Kingma & Ba, 2015: the
first bias-corrected step is lr * g / (|g| + eps) per coordinate).
"""
import numpy as np


def layout(dims):
    """[(w_offset, b_offset, dout, din)] per layer and the total parameter count"""
    out, off = [], 0
    for din, dout in zip(dims[:-1], dims[1:]):
        out.append((off, off + din * dout, dout, din))
        off += din * dout + dout
    return out, off


def init_params(dims, seed):
    """Glorot-uniform weights, zero biases (float32, the device's dtype)"""
    rng = np.random.default_rng(seed)
    lay, n = layout(dims)
    p = np.zeros(n, np.float32)
    for wo, bo, dout, din in lay:
        lim = np.sqrt(6.0 / (din + dout))
        p[wo:bo] = rng.uniform(-lim, lim, size=din * dout).astype(np.float32)
    return p


def forward(p, dims, x):
    """list of layer outputs [x, h_1, ..., y] (float64)"""
    lay, _ = layout(dims)
    p = np.asarray(p, np.float64)
    hs = [np.asarray(x, np.float64)]
    for l, (wo, bo, dout, din) in enumerate(lay):
        W = p[wo:bo].reshape(dout, din)
        z = hs[-1] @ W.T + p[bo:bo + dout]
        hs.append(np.tanh(z) if l + 1 < len(lay) else z)
    return hs


def loss_grad(p, dims, x, t, w=None):
    """(loss, flat gradient) of the weighted squared loss"""
    lay, n = layout(dims)
    p = np.asarray(p, np.float64)
    hs = forward(p, dims, x)
    B = hs[0].shape[0]
    w = np.ones(dims[-1]) if w is None else np.asarray(w, np.float64)
    e = hs[-1] - np.asarray(t, np.float64)
    loss = 0.5 * np.sum(w * e * e) / B
    g = np.zeros(n)
    d = w * e / B
    for l in range(len(lay) - 1, -1, -1):
        wo, bo, dout, din = lay[l]
        g[wo:bo] = (d.T @ hs[l]).reshape(-1)
        g[bo:bo + dout] = d.sum(axis=0)
        if l > 0:
            W = p[wo:bo].reshape(dout, din)
            d = (d @ W) * (1.0 - hs[l] ** 2)
    return loss, g


def adam_step(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """one Adam update (step counts from 1); returns (p, m, v)"""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
