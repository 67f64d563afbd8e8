"""The surrogate U-Net on the GPU (ludwig_amd/surrogate.py, every pass on
the library's kernels) against the FP64 reference network (tests/
unet_ref.py, itself pinned by finite differences) with the same BF16
rounding points: the forward output and every layer's gradients.  Then
end-to-end training: Adam on the GPU learns a fixed smooth periodic map."""

import ctypes

import numpy as np
import pytest

import ludwig_amd as L
from bf16util import bf16_bits, bf16_round
from ludwig_amd.surrogate import LAYERS, UNet, loss_and_grad
from unet_ref import RefUNet

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_unet_forward_and_gradients_match_reference():
    shape, batch = (16, 16, 8), 2
    n = batch * int(np.prod(shape))
    rng = np.random.default_rng(11)
    net = UNet(seed=5, out_scale=1.0)
    p = net.params().astype(np.float64)
    x = rng.standard_normal((n, 16))
    x[:, 7:] = 0.0
    xb = bf16_bits(x.ravel())
    dx = L.DeviceArray.from_host(xb, np.uint16)
    print(f"unet forward/backward batch={batch} shape={shape}", flush=True)
    y = net.forward(dx, batch, shape)
    yg = y.download().astype(np.float64)
    ref = RefUNet(p, quantise=True)
    yr = ref.forward(bf16_round(x.ravel()), batch, shape)
    assert rel(yg, yr) < 5e-3, rel(yg, yr)

    t = (rng.standard_normal((n, 4)) * 0.5).astype(np.float32)
    dt = L.DeviceArray.from_host(t, np.float32)
    dyb, dyf = L.DeviceArray(n * 16, np.uint16), L.DeviceArray(n * 16, np.float32)
    loss = loss_and_grad(y, dt, n, dyb, dyf)
    e = yg.reshape(n, 16)[:, :4] - t
    assert loss == pytest.approx(np.mean(e * e), rel=1e-5)
    net.backward(dyb, dyf)
    gg = net.grads().astype(np.float64)
    gr = ref.backward(dyf.download().astype(np.float64))
    o = 0
    for name, ci, co in LAYERS:
        nw = 27 * co * ci
        w_err = rel(gg[o:o + nw], gr[o:o + nw])
        b_err = rel(gg[o + nw:o + nw + co], gr[o + nw:o + nw + co])
        assert w_err < 3e-2 and b_err < 3e-2, (name, w_err, b_err)
        o += nw + co


def smooth_field(rng, shape, nmodes=6):
    """A sum of a few random periodic Fourier modes, scaled to unit RMS."""
    nx, ny, nz = shape
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    f = np.zeros(shape)
    for _ in range(nmodes):
        k = rng.integers(-2, 3, 3)
        ph = rng.uniform(0, 2 * np.pi)
        f += np.cos(2 * np.pi * (k[0] * X / nx + k[1] * Y / ny + k[2] * Z / nz) + ph)
    return (f / max(np.sqrt(np.mean(f * f)), 1e-12)).ravel()


def smooth_target(phi, shape):
    """A fixed periodic map: the 7-point average of phi plus 0.2 phi^3."""
    a = phi.reshape(shape)
    avg = (a + sum(np.roll(a, s, axis=k) for k in range(3) for s in (1, -1))) / 7
    return (avg + 0.2 * a ** 3).ravel()


def test_training_learns_a_smooth_map():
    shape, batch = (16, 16, 16), 4
    ns = int(np.prod(shape))
    n = batch * ns
    rng = np.random.default_rng(12)
    net = UNet(seed=7)
    xs, ts = [], []
    for k in range(8):
        phi = smooth_field(rng, shape)
        x = np.zeros((ns, 16))
        x[:, 0] = phi
        xs.append(x)
        t = np.zeros((ns, 4))
        t[:, 0] = smooth_target(phi, shape)
        ts.append(t)
    dyb, dyf = L.DeviceArray(n * 16, np.uint16), L.DeviceArray(n * 16, np.float32)
    losses = []
    for it in range(300):
        idx = rng.choice(len(xs), batch, replace=False)
        dx = L.DeviceArray.from_host(bf16_bits(np.concatenate([xs[i] for i in idx]).ravel()), np.uint16)
        dt = L.DeviceArray.from_host(np.concatenate([ts[i] for i in idx]).astype(np.float32), np.float32)
        y = net.forward(dx, batch, shape)
        losses.append(loss_and_grad(y, dt, n, dyb, dyf))
        net.backward(dyb, dyf)
        net.adam(1e-3 * min(1.0, (it + 1) / 20))
        dx.free(); dt.free()
    print(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
    assert np.isfinite(losses).all()
    assert np.mean(losses[-10:]) < 0.2 * np.mean(losses[:5])
