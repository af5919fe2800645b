"""The device warm-start network (qh_mlp.hip) against the CPU oracle.

Tolerances: the device computes in FP32, the oracle in FP64 on the same FP32
parameters and inputs; forward outputs and loss agree to 1e-5 relative,
gradients to 1e-4 of the largest gradient entry, one Adam step to 1e-6
absolute (lr 1e-3)."""
import os
import sys

import numpy as np
import pytest

import quac_amd as qa

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402

pytestmark = pytest.mark.gpu

SHAPES = [([10, 256, 256, 256, 64], 37), ([3, 5, 2], 1), ([7, 70, 130, 9], 129)]


def _setup(dims, B, seed):
    rng = np.random.default_rng(seed)
    p = mlp_ref.init_params(dims, seed)
    x = rng.normal(size=(B, dims[0])).astype(np.float32)
    t = rng.normal(size=(B, dims[-1])).astype(np.float32)
    net = qa.MLP(dims, max(B, 4))
    net.set_params(p)
    xd, td = qa.DeviceF32(x.size), qa.DeviceF32(t.size)
    xd.set(x)
    td.set(t)
    return net, p, x, t, xd, td


@pytest.mark.parametrize("dims,B", SHAPES)
def test_forward_matches_oracle(dims, B):
    net, p, x, t, xd, td = _setup(dims, B, 5)
    yd = qa.DeviceF32(B * dims[-1])
    net.forward(xd, B, yd)
    y = yd.get().reshape(B, dims[-1])
    ref = mlp_ref.forward(p, dims, x)[-1]
    assert np.abs(y - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())
    net.destroy()


@pytest.mark.parametrize("dims,B", SHAPES)
def test_loss_and_gradient_match_oracle(dims, B):
    net, p, x, t, xd, td = _setup(dims, B, 6)
    w = np.linspace(0.5, 2.0, dims[-1]).astype(np.float32)
    net.set_output_weights(w)
    loss, g = net.loss_grad(xd, td, B)
    lref, gref = mlp_ref.loss_grad(p, dims, x, t, w)
    assert abs(loss - lref) <= 1e-5 * abs(lref)
    assert np.abs(g - gref).max() <= 1e-4 * np.abs(gref).max()
    net.destroy()


def test_adam_step_matches_oracle():
    dims, B = [10, 64, 64, 16], 50
    net, p, x, t, xd, td = _setup(dims, B, 7)
    net.train_step(xd, td, B, lr=1e-3)
    _, g = mlp_ref.loss_grad(p, dims, x, t)
    pref, _, _ = mlp_ref.adam_step(p.astype(np.float64), g, np.zeros_like(g), np.zeros_like(g), 1, lr=1e-3)
    assert np.abs(net.get_params() - pref).max() < 1e-6
    net.destroy()


def test_training_fits_a_smooth_map():
    """a 10 -> 64 map that the network can represent; the loss falls 100x"""
    rng = np.random.default_rng(8)
    dims, B = [10, 128, 128, 64], 512
    x = rng.uniform(-1, 1, size=(B, 10)).astype(np.float32)
    A = rng.normal(size=(10, 64)) / 3
    t = np.tanh(x @ A).astype(np.float32)
    net = qa.MLP(dims, B)
    net.set_params(mlp_ref.init_params(dims, 9))
    xd, td = qa.DeviceF32(x.size), qa.DeviceF32(t.size)
    xd.set(x)
    td.set(t)
    first = net.train_step(xd, td, B, lr=3e-3, want_loss=True)
    for _ in range(600):
        net.train_step(xd, td, B, lr=3e-3)
    last, _ = net.loss_grad(xd, td, B)
    assert last < first / 100, (first, last)
    net.destroy()
