"""The FP64 reference U-Net (tests/unet_ref.py) pinned by central finite
differences of its own loss: its backward pass, composed from the
convolution checker's adjoints and numpy, is the gradient of its forward
pass.  Without BF16 rounding the loss is smooth, so the differences agree
to ~1e-7."""

import numpy as np
import pytest

from ludwig_amd.surrogate import LAYERS
from unet_ref import RefUNet


def make_params(seed):
    rng = np.random.default_rng(seed)
    parts = []
    for name, ci, co in LAYERS:
        parts.append(rng.standard_normal(27 * co * ci) * np.sqrt(2.0 / (27 * ci)))
        parts.append(rng.standard_normal(co) * 0.1)
    return np.concatenate(parts)


def test_reference_gradient_by_finite_differences():
    shape, batch = (8, 8, 8), 1
    n = batch * 512
    rng = np.random.default_rng(3)
    p = make_params(4)
    x = rng.standard_normal((n, 16))
    x[:, 7:] = 0.0
    x = x.ravel()
    t = rng.standard_normal((n, 4)) * 0.1

    def loss(params):
        net = RefUNet(params, quantise=False)
        y = net.forward(x, batch, shape).reshape(n, 16)
        e = y[:, :4] - t
        return float(np.mean(e * e)), net, y

    l0, net, y = loss(p)
    dy = np.zeros((n, 16))
    dy[:, :4] = 2 * (y[:, :4] - t) / (4 * n)
    g = net.backward(dy.ravel())
    # parameters from several layers, weights and biases
    offs, o = {}, 0
    for name, ci, co in LAYERS:
        offs[name] = (o, ci, co)
        o += 27 * co * ci + co
    picks = []
    for name in ("e0a", "e1b", "b1", "d2", "d0", "out"):
        o, ci, co = offs[name]
        picks += [o + int(rng.integers(27 * co * ci)), o + 27 * co * ci + int(rng.integers(co))]
    h = 1e-5
    for i in picks:
        e = np.zeros_like(p); e[i] = h
        fd = (loss(p + e)[0] - loss(p - e)[0]) / (2 * h)
        assert fd == pytest.approx(g[i], rel=2e-5, abs=1e-9), i
