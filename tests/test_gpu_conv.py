"""The MFMA convolutions and the network's elementwise kernels through the
C ABI, against the FP64 CPU checker (oracle/conv_oracle.c) on the same
BF16-rounded inputs.

Tolerance: products of two BF16 values are exact in FP32, so the only error
is FP32 accumulation of K terms: |gpu - exact| <= 2 K 2^-24 sum |a b| per
output element (K = 27 cin for the data passes, the number of positions for
the weight gradient), with sum |a b| from the checker applied to |a|, |b|.
"""

import ctypes

import numpy as np
import pytest

import ludwig_amd as L
import oracle as O
from bf16util import bf16_bits, bf16_round, bf16_value

pytestmark = pytest.mark.gpu

SHAPE = (6, 8, 10)
EPS32 = 2.0 ** -24


def dev(a, dtype):
    return L.DeviceArray.from_host(np.ascontiguousarray(a), dtype)


def vp(d):
    return ctypes.c_void_p(d.ptr)


def within(got, want, bound, what):
    err = np.abs(got - want)
    bad = np.flatnonzero(err > bound)
    if bad.size:
        i = bad[0]
        pytest.fail(f"{what}: {bad.size}/{got.size} outside bound, first {i}: {got[i]} vs {want[i]} "
                    f"(bound {bound[i]:.3e})")


# lattices below 4 x 4 x 16 take the fragment-direct kernel, the rest the
# LDS-tiled one (ragged tiles and 16-channel chunks included)
@pytest.mark.parametrize("cin,cout,batch,shape", [
    (16, 16, 1, SHAPE), (32, 48, 2, SHAPE), (48, 64, 1, SHAPE), (16, 96, 2, SHAPE),
    (32, 160, 1, SHAPE), (160, 32, 1, SHAPE),
    (32, 96, 2, (8, 4, 16)), (48, 64, 1, (5, 6, 20)), (16, 16, 1, (4, 4, 32)),
    (80, 112, 1, (6, 5, 17)), (160, 64, 1, (4, 8, 16)), (48, 32, 2, (8, 8, 16)),
    (32, 16, 1, (4, 12, 16))])
def test_conv_fwd_bwd_against_oracle(cin, cout, batch, shape):
    SHAPE = shape
    lib = L.lib()
    rng = np.random.default_rng(cin * 1000 + cout)
    n = batch * int(np.prod(SHAPE))
    print(f"conv cin={cin} cout={cout} batch={batch} shape={SHAPE}", flush=True)
    x = bf16_round(rng.standard_normal(n * cin))
    w32 = (rng.standard_normal(27 * cout * cin) / np.sqrt(27 * cin)).astype(np.float32)
    w = bf16_round(w32)
    bias = rng.standard_normal(cout).astype(np.float32)
    g = bf16_round(rng.standard_normal(n * cout))

    dx_in, dw_in, dg = dev(bf16_bits(x), np.uint16), dev(w32, np.float32), dev(bf16_bits(g), np.uint16)
    wf, wd = L.DeviceArray(w.size, np.uint16), L.DeviceArray(w.size, np.uint16)
    db = dev(bias, np.float32)
    y, dxo, dwo = L.DeviceArray(n * cout, np.float32), L.DeviceArray(n * cin, np.float32), \
        L.DeviceArray(w.size, np.float32)
    assert lib.ldw_conv3d_weights_prep(vp(dw_in), vp(wf), vp(wd), cin, cout) == 0
    assert np.array_equal(wf.download(), bf16_bits(w32))
    nx, ny, nz = SHAPE
    assert lib.ldw_conv3d_fwd(vp(dx_in), vp(wf), vp(db), vp(y), batch, nx, ny, nz, cin, cout, 0) == 0
    assert lib.ldw_conv3d_bwd_data(vp(dg), vp(wd), vp(dxo), batch, nx, ny, nz, cin, cout, 0) == 0
    assert lib.ldw_conv3d_bwd_weight(vp(dx_in), vp(dg), vp(dwo), batch, nx, ny, nz, cin, cout, 0) == 0

    yo = O.conv3_fwd(x, w, bias.astype(np.float64), SHAPE, cin, cout, batch)
    ya = O.conv3_fwd(np.abs(x), np.abs(w), np.abs(bias.astype(np.float64)), SHAPE, cin, cout, batch)
    within(y.download().astype(np.float64), yo, 2 * 27 * cin * EPS32 * ya + 1e-30, "forward")
    dxr = O.conv3_bwd_data(g, w, SHAPE, cin, cout, batch)
    dxa = O.conv3_bwd_data(np.abs(g), np.abs(w), SHAPE, cin, cout, batch)
    within(dxo.download().astype(np.float64), dxr, 2 * 27 * cout * EPS32 * dxa + 1e-30, "backward data")
    dwr = O.conv3_bwd_weight(x, g, SHAPE, cin, cout, batch)
    dwa = O.conv3_bwd_weight(np.abs(x), np.abs(g), SHAPE, cin, cout, batch)
    within(dwo.download().astype(np.float64), dwr, 2 * n * EPS32 * dwa + 1e-30, "backward weight")
    # SiLU fused into the epilogue: the same pre-activation and, bit for bit,
    # what the standalone SiLU kernel makes of it
    z2, a2, a1 = L.DeviceArray(n * cout, np.float32), L.DeviceArray(n * cout, np.uint16), \
        L.DeviceArray(n * cout, np.uint16)
    assert lib.ldw_conv3d_fwd_silu(vp(dx_in), vp(wf), vp(db), vp(z2), vp(a2), batch, nx, ny, nz, cin, cout) == 0
    assert np.array_equal(z2.download(), y.download())
    assert lib.ldw_nn_silu_fwd(vp(y), vp(a1), n * cout) == 0
    assert np.array_equal(a2.download(), a1.download())
    assert lib.ldw_conv3d_fwd_silu(vp(dx_in), vp(wf), vp(db), None, vp(a2), batch, nx, ny, nz, cin, cout) == 0
    assert np.array_equal(a2.download(), a1.download())
    # accumulate = 1 adds to what is there
    assert lib.ldw_conv3d_fwd(vp(dx_in), vp(wf), vp(db), vp(y), batch, nx, ny, nz, cin, cout, 1) == 0
    within(y.download().astype(np.float64), 2 * yo, 4 * 27 * cin * EPS32 * ya + 1e-30, "forward x2")


def test_pool_up_concat_split_silu_exact():
    lib = L.lib()
    rng = np.random.default_rng(1)
    b, nx, ny, nz, c = 2, 4, 6, 8, 16
    n = b * nx * ny * nz
    a = bf16_round(rng.standard_normal(n * c))
    da = dev(bf16_bits(a), np.uint16)
    po = L.DeviceArray(n // 8 * c, np.uint16)
    assert lib.ldw_nn_pool2_fwd(vp(da), vp(po), b, nx, ny, nz, c) == 0
    a5 = a.reshape(b, nx // 2, 2, ny // 2, 2, nz // 2, 2, c)
    want = bf16_round(a5.mean(axis=(2, 4, 6)).astype(np.float32).ravel())
    got = bf16_value(po.download())
    # the kernel sums 8 BF16 values in FP32 (exact here) then scales by 1/8
    assert np.array_equal(got, want)
    up = L.DeviceArray(n * c, np.uint16)
    assert lib.ldw_nn_up2_fwd(vp(po), vp(up), b, nx // 2, ny // 2, nz // 2, c) == 0
    u = bf16_value(up.download()).reshape(b, nx, ny, nz, c)
    pc = got.reshape(b, nx // 2, ny // 2, nz // 2, c)
    assert np.array_equal(u, pc.repeat(2, 1).repeat(2, 2).repeat(2, 3))
    # backward of up: sum of children; of pool: parent / 8
    g = rng.standard_normal(n * c).astype(np.float32)
    dg, dc = dev(g, np.float32), L.DeviceArray(n // 8 * c, np.float32)
    assert lib.ldw_nn_up2_bwd(vp(dg), vp(dc), b, nx // 2, ny // 2, nz // 2, c) == 0
    g5 = g.astype(np.float64).reshape(b, nx // 2, 2, ny // 2, 2, nz // 2, 2, c)
    assert np.allclose(dc.download(), g5.sum(axis=(2, 4, 6)).ravel(), rtol=1e-6, atol=1e-6)
    df = L.DeviceArray(n * c, np.float32)
    assert lib.ldw_nn_pool2_bwd(vp(dc), vp(df), b, nx, ny, nz, c, 0) == 0
    dcv = dc.download().reshape(b, nx // 2, ny // 2, nz // 2, c)
    assert np.array_equal(df.download().reshape(b, nx, ny, nz, c),
                          (np.float32(0.125) * dcv).repeat(2, 1).repeat(2, 2).repeat(2, 3))
    # nearest upsampling and concatenation in one pass
    skip = bf16_bits(rng.standard_normal(n * 32))
    dsk, uc = dev(skip, np.uint16), L.DeviceArray(n * (c + 32), np.uint16)
    assert lib.ldw_nn_up2_concat(vp(po), c, vp(dsk), 32, vp(uc), b, nx, ny, nz) == 0
    want_uc = np.concatenate([bf16_bits(u.ravel()).reshape(n, c), skip.reshape(n, 32)], axis=1)
    assert np.array_equal(uc.download().reshape(n, c + 32), want_uc)
    # concat / split
    c2 = 32
    a2 = bf16_bits(rng.standard_normal(n * c2))
    d2 = dev(a2, np.uint16)
    cat = L.DeviceArray(n * (c + c2), np.uint16)
    assert lib.ldw_nn_concat(vp(da), c, vp(d2), c2, vp(cat), n) == 0
    assert np.array_equal(cat.download().reshape(n, c + c2),
                          np.concatenate([bf16_bits(a).reshape(n, c), a2.reshape(n, c2)], axis=1))
    gg = rng.standard_normal(n * (c + c2)).astype(np.float32)
    dgg, sa, sb = dev(gg, np.float32), L.DeviceArray(n * c, np.float32), L.DeviceArray(n * c2, np.float32)
    assert lib.ldw_nn_split(vp(dgg), c, c2, vp(sa), vp(sb), n, 0, 0) == 0
    assert np.array_equal(sa.download(), gg.reshape(n, -1)[:, :c].ravel())
    assert np.array_equal(sb.download(), gg.reshape(n, -1)[:, c:].ravel())
    # SiLU and its derivative against float64
    z = rng.standard_normal(n * c).astype(np.float32) * 3
    dz, act = dev(z, np.float32), L.DeviceArray(n * c, np.uint16)
    assert lib.ldw_nn_silu_fwd(vp(dz), vp(act), n * c) == 0
    zz = z.astype(np.float64)
    sig = 1 / (1 + np.exp(-zz))
    assert np.allclose(bf16_value(act.download()), zz * sig, rtol=2 ** -7, atol=1e-6)
    dgr = dev(g.astype(np.float32), np.float32)
    dzo, dzb = L.DeviceArray(n * c, np.float32), L.DeviceArray(n * c, np.uint16)
    assert lib.ldw_nn_silu_bwd(vp(dz), vp(dgr), vp(dzo), vp(dzb), n * c) == 0
    assert np.allclose(dzo.download(), g * sig * (1 + zz * (1 - sig)), rtol=1e-5, atol=1e-6)
    # bias gradient: per-channel sums
    dbias = L.DeviceArray(c, np.float32)
    assert lib.ldw_nn_bias_grad(vp(dgr), vp(dbias), n, c, 0) == 0
    assert np.allclose(dbias.download(), g.astype(np.float64).reshape(n, c).sum(0), rtol=1e-5, atol=1e-4)


def test_mse_and_adam():
    lib = L.lib()
    rng = np.random.default_rng(2)
    npos, cy = 1000, 16
    y = rng.standard_normal(npos * cy).astype(np.float32)
    t = rng.standard_normal(npos * 4).astype(np.float32)
    dy_, dt_, dg = dev(y, np.float32), dev(t, np.float32), L.DeviceArray(npos * cy, np.uint16)
    loss = ctypes.c_double()
    dgf = L.DeviceArray(npos * cy, np.float32)
    assert lib.ldw_nn_mse(vp(dy_), cy, vp(dt_), npos, vp(dg), vp(dgf), None, ctypes.byref(loss)) == 0
    e = y.reshape(npos, cy)[:, :4].astype(np.float64) - t.reshape(npos, 4)
    assert loss.value == pytest.approx(np.mean(e * e), rel=1e-6)
    grad = np.zeros((npos, cy))
    grad[:, :4] = 2 * e / (4 * npos)
    assert np.allclose(bf16_value(dg.download()), grad.ravel(), rtol=2 ** -8, atol=1e-12)
    assert np.allclose(dgf.download(), grad.ravel(), rtol=1e-6, atol=1e-12)
    n = 777
    w, g = rng.standard_normal(n).astype(np.float32), rng.standard_normal(n).astype(np.float32)
    m, v = np.zeros(n, np.float32), np.zeros(n, np.float32)
    dw, dgr, dm, dv = (dev(a, np.float32) for a in (w, g, m, v))
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    wref, mref, vref = w.astype(np.float64), m.astype(np.float64), v.astype(np.float64)
    for step in (1, 2, 3):
        assert lib.ldw_nn_adam(vp(dw), vp(dgr), vp(dm), vp(dv), n, lr, b1, b2, eps, 0.0, 1.0, step) == 0
        mref = b1 * mref + (1 - b1) * g
        vref = b2 * vref + (1 - b2) * g.astype(np.float64) ** 2
        mh, vh = mref / (1 - b1 ** step), vref / (1 - b2 ** step)
        wref = wref - lr * mh / (np.sqrt(vh) + eps)
    assert np.allclose(dw.download(), wref, rtol=1e-5, atol=1e-6)
    # weighted loss and the sum of squares used for gradient clipping
    wts = (ctypes.c_float * 4)(1.0, 0.5, 0.25, 0.0)
    assert lib.ldw_nn_mse(vp(dy_), cy, vp(dt_), npos, vp(dg), vp(dgf), wts, ctypes.byref(loss)) == 0
    wv = np.array([1.0, 0.5, 0.25, 0.0])
    assert loss.value == pytest.approx(np.mean(wv * e * e), rel=1e-6)
    assert np.allclose(dgf.download().reshape(npos, cy)[:, :4], 2 * wv * e / (4 * npos), rtol=1e-6, atol=1e-12)
    ss = ctypes.c_double()
    assert lib.ldw_nn_sumsq(vp(dgr), n, ctypes.byref(ss)) == 0
    assert ss.value == pytest.approx(float(np.sum(g.astype(np.float64) ** 2)), rel=1e-12)


def test_sum_shift_and_restore_sums():
    """The jump's conservation step: plane sums of the new state are set to
    those of the old one by a uniform shift of each plane."""
    lib = L.lib()
    rng = np.random.default_rng(5)
    ns, nplane = 4099, 4
    now = rng.standard_normal(nplane * ns).astype(np.float32)
    out = (now + 0.01 + 0.1 * rng.standard_normal(nplane * ns)).astype(np.float32)
    dnow, dout = dev(now, np.float32), dev(out, np.float32)
    s = ctypes.c_double()
    assert lib.ldw_nn_sum(vp(dnow), ns, ctypes.byref(s)) == 0
    assert s.value == pytest.approx(float(np.sum(now[:ns].astype(np.float64))), rel=1e-12, abs=1e-9)
    assert lib.ldw_nn_restore_sums(vp(dnow), vp(dout), ns, nplane) == 0
    got = dout.download().reshape(nplane, ns).astype(np.float64)
    want_shift = (now.reshape(nplane, ns).astype(np.float64).sum(1)
                  - out.reshape(nplane, ns).astype(np.float64).sum(1)) / ns
    assert np.allclose(got - out.reshape(nplane, ns), want_shift[:, None], atol=2e-7)
    assert np.allclose(got.sum(1), now.reshape(nplane, ns).astype(np.float64).sum(1), atol=ns * 2e-7)
    assert lib.ldw_nn_shift(vp(dnow), ns, 0.5) == 0
    assert np.allclose(dnow.download()[:ns], now[:ns] + 0.5, atol=1e-6)


def np_cubic(state, n, perm, flip):
    """The cubic map of ldw_nn_cubic in numpy: state (4, n, n, n)."""
    def move(a):
        b = np.transpose(a, perm)
        for ax in range(3):
            if flip[ax]:
                b = np.flip(b, axis=ax)
        return b
    out = np.empty_like(state)
    out[0] = move(state[0])
    for ax in range(3):
        out[1 + ax] = (-1.0 if flip[ax] else 1.0) * move(state[1 + perm[ax]])
    return out


def test_cubic_maps_against_numpy():
    """ldw_nn_cubic for every element of the cubic group (the training
    augmentation) against the same map in numpy, bit for bit."""
    from ludwig_amd.surrogate import CUBIC_GROUP, cubic
    n = 6
    rng = np.random.default_rng(8)
    s = rng.standard_normal((4, n, n, n)).astype(np.float32)
    src = dev(s.ravel(), np.float32)
    dst = L.DeviceArray(4 * n ** 3, np.float32)
    assert len(CUBIC_GROUP) == 48 and len(set(CUBIC_GROUP)) == 48
    for perm, flip in CUBIC_GROUP:
        cubic(src, dst, n, perm, flip)
        got = dst.download().reshape(4, n, n, n)
        assert np.array_equal(got, np_cubic(s, n, perm, flip)), (perm, flip)
    lib = L.lib()
    lib.ldw_error_mode_set(0)
    try:
        bad = (ctypes.c_int * 3)(0, 0, 2)
        ok = (ctypes.c_int * 3)(0, 0, 0)
        assert lib.ldw_nn_cubic(vp(src), vp(dst), n, bad, ok) == -1
    finally:
        lib.ldw_error_mode_set(1)
    src.free()
    dst.free()


@pytest.mark.parametrize("perm,flip", [((0, 1, 2), (1, 0, 0)), ((2, 0, 1), (0, 1, 1)), ((1, 0, 2), (0, 0, 0))])
def test_binary_fluid_commutes_with_cubic_maps(perm, flip):
    """What makes the augmentation sound: the lattice Boltzmann binary
    fluid run from a mapped phi equals the mapped run (FP64, 16^3, 12
    steps), to rounding -- the stencils sum their terms in a fixed order,
    so a map changes the order of some additions."""
    n, steps = 16, 12
    p = dict(a=-0.0625, b=0.0625, kappa=0.04, mobility=0.25, tau=1.0, fatal=True)
    b = L.BinaryFluid((n, n, n), **p)
    b.init_spinodal(77, 0.0, 0.1)
    phi0 = b.phi_field()
    b.step(steps)
    rho, u, _ = b.hydro_fields()
    s = np.concatenate([b.phi_field()[None], u.reshape(3, -1)]).reshape(4, n, n, n)
    b.close()
    s0 = np.concatenate([phi0[None], np.zeros((3, n ** 3))]).reshape(4, n, n, n)
    c = L.BinaryFluid((n, n, n), **p)
    c.init_spinodal(77, 0.0, 0.1)
    c.set_phi(np_cubic(s0, n, perm, flip)[0].ravel())
    c.step(steps)
    _, u2, _ = c.hydro_fields()
    got = np.concatenate([c.phi_field()[None], u2.reshape(3, -1)]).reshape(4, n, n, n)
    c.close()
    want = np_cubic(s, n, perm, flip)
    assert np.max(np.abs(got[0] - want[0])) < 1e-12
    assert np.max(np.abs(got[1:] - want[1:])) < 1e-13
    # and the map is not the identity on this state
    assert np.max(np.abs(want[0] - s[0])) > 0.01 or perm == (0, 1, 2) and not any(flip)


@pytest.mark.parametrize("c,cy", [(16, 16), (12, 6)])
def test_encode_decode_against_numpy(c, cy):
    """The network's input encoding and the jump's decoding, bit for bit
    against float32 numpy: per-site kernels (channel counts that allow
    16-byte vectors) and element kernels (those that do not)."""
    rng = np.random.default_rng(c * 100 + cy)
    ns = 1000
    state = rng.standard_normal(4 * ns).astype(np.float32)
    norm = np.array([0.1, 1.7, 25.0, 0.3, -0.2, 0.9], np.float32)
    print(f"encode/decode c={c} cy={cy}", flush=True)
    ds, dx = dev(state, np.float32), L.DeviceArray(ns * c, np.uint16)
    lib = L.lib()
    assert lib.ldw_nn_encode(vp(ds), ns, norm.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), vp(dx), c) == 0
    want = np.zeros((ns, c), np.float32)
    want[:, 0] = (state[:ns] - norm[0]) * norm[1]
    for k in range(1, 4):
        want[:, k] = state[k * ns:(k + 1) * ns] * norm[2]
    want[:, 4:7] = norm[3:6]
    assert np.array_equal(dx.download().reshape(ns, c), bf16_bits(want)), "encode"
    y = rng.standard_normal(ns * cy).astype(np.float32)
    dy, do = dev(y, np.float32), L.DeviceArray(4 * ns, np.float32)
    sc = np.float32(0.05), np.float32(0.002)
    assert lib.ldw_nn_decode(vp(ds), vp(dy), cy, ns, vp(do), sc[0], sc[1], 1.5, 0.1) == 0
    got = do.download().reshape(4, ns)
    yy = y.reshape(ns, cy)
    for k in range(4):
        m = np.float32(1.5 if k == 0 else 0.1)
        v = state[k * ns:(k + 1) * ns] + yy[:, k] / (sc[0] if k == 0 else sc[1])
        assert np.array_equal(got[k], np.minimum(np.maximum(v, -m), m)), f"decode channel {k}"
    for d in (ds, dx, dy, do):
        d.free()
