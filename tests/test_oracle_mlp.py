"""The CPU restatement of the warm-start network, pinned by mathematics:
central finite differences for the gradient, the closed-form first Adam step,
and the reduction of a one-layer network to a linear least-squares model."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402


def test_gradient_matches_central_differences():
    dims = [5, 7, 6, 3]
    rng = np.random.default_rng(0)
    p = mlp_ref.init_params(dims, 1).astype(np.float64) + 0.05 * rng.normal(size=mlp_ref.layout(dims)[1])
    x = rng.normal(size=(4, 5))
    t = rng.normal(size=(4, 3))
    w = rng.uniform(0.5, 2.0, size=3)
    loss, g = mlp_ref.loss_grad(p, dims, x, t, w)
    h = 1e-6
    fd = np.zeros_like(g)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (mlp_ref.loss_grad(p + e, dims, x, t, w)[0] - mlp_ref.loss_grad(p - e, dims, x, t, w)[0]) / (2 * h)
    assert np.abs(fd - g).max() < 1e-7 * max(1.0, np.abs(g).max())
    assert loss > 0


def test_first_adam_step_is_lr_times_sign():
    rng = np.random.default_rng(2)
    g = rng.normal(size=50)
    p0 = rng.normal(size=50)
    p1, m, v = mlp_ref.adam_step(p0, g, np.zeros(50), np.zeros(50), 1, lr=1e-2, eps=1e-12)
    assert np.abs((p0 - p1) - 1e-2 * np.sign(g)).max() < 1e-9


def test_single_linear_layer_is_least_squares():
    """one Linear layer, unit weights: the gradient vanishes at the least-squares solution"""
    rng = np.random.default_rng(3)
    x = rng.normal(size=(40, 4))
    t = rng.normal(size=(40, 2))
    X1 = np.hstack([x, np.ones((40, 1))])
    sol = np.linalg.lstsq(X1, t, rcond=None)[0]          # (5, 2)
    p = np.concatenate([sol[:4].T.reshape(-1), sol[4]])   # W (2 x 4) row-major, then b
    _, g = mlp_ref.loss_grad(p, [4, 2], x, t)
    assert np.abs(g).max() < 1e-12
