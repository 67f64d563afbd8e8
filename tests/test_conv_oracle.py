"""The FP64 convolution checker (oracle/conv_oracle.c) against what fixes
it independently of its own code:

* the adjoint relations: bwd_data(g) and bwd_weight(x, g) are the gradients
  of L = sum g . conv(x, w), checked by central finite differences (L is
  linear in x and in w, so the differences are exact up to rounding);
* closed forms: the centre tap as identity gives y = x, a single off-centre
  tap gives the periodic shift, bias alone gives a constant;
* torch's CPU conv3d with circular padding (an independent implementation),
  after mapping w[t][co][ci] -> weight[co][ci][dx+1][dy+1][dz+1].
"""

import numpy as np
import pytest

import oracle as O

SHAPE = (3, 4, 5)


def rand(rng, *n):
    return rng.standard_normal(int(np.prod(n)))


@pytest.mark.parametrize("cin,cout,batch", [(3, 2, 1), (2, 4, 2)])
def test_adjoints_by_finite_differences(cin, cout, batch):
    rng = np.random.default_rng(5)
    n = batch * int(np.prod(SHAPE))
    x = rand(rng, n, cin)
    w = rand(rng, 27, cout, cin)
    g = rand(rng, n, cout)
    L = lambda xx, ww: float(g @ O.conv3_fwd(xx, ww, None, SHAPE, cin, cout, batch))
    dx = O.conv3_bwd_data(g, w, SHAPE, cin, cout, batch)
    dw = O.conv3_bwd_weight(x, g, SHAPE, cin, cout, batch)
    h = 0.5
    for i in rng.choice(x.size, 12, replace=False):
        e = np.zeros_like(x); e[i] = h
        fd = (L(x + e, w) - L(x - e, w)) / (2 * h)
        assert fd == pytest.approx(dx[i], rel=1e-10, abs=1e-10)
    for i in rng.choice(w.size, 12, replace=False):
        e = np.zeros_like(w); e[i] = h
        fd = (L(x, w + e) - L(x, w - e)) / (2 * h)
        assert fd == pytest.approx(dw[i], rel=1e-10, abs=1e-10)


def test_closed_forms():
    rng = np.random.default_rng(6)
    nx, ny, nz = SHAPE
    c = 3
    x = rand(rng, nx * ny * nz, c)
    w = np.zeros((27, c, c))
    w[13] = np.eye(c)                     # centre tap (0, 0, 0)
    assert np.array_equal(O.conv3_fwd(x, w.ravel(), None, SHAPE, c, c), x)
    # tap (dx, dy, dz) = (1, -1, 0): y(p) = x(p + o)
    t = 9 * 2 + 3 * 0 + 1
    w = np.zeros((27, c, c)); w[t] = np.eye(c)
    y = O.conv3_fwd(x, w.ravel(), None, SHAPE, c, c).reshape(nx, ny, nz, c)
    xs = x.reshape(nx, ny, nz, c)
    assert np.array_equal(y, np.roll(xs, shift=(-1, 1, 0), axis=(0, 1, 2)))
    b = np.array([0.5, -2.0])
    y = O.conv3_fwd(x, np.zeros(27 * 2 * c), b, SHAPE, c, 2).reshape(-1, 2)
    assert np.all(y == b)


def test_against_torch_circular_conv3d():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(7)
    nx, ny, nz = SHAPE
    cin, cout, batch = 4, 3, 2
    x = rand(rng, batch, nx, ny, nz, cin)
    w = rand(rng, 27, cout, cin)
    bias = rand(rng, cout)
    y = O.conv3_fwd(x, w, bias, SHAPE, cin, cout, batch).reshape(batch, nx, ny, nz, cout)
    xt = torch.from_numpy(x.reshape(batch, nx, ny, nz, cin)).permute(0, 4, 1, 2, 3)
    wt = torch.from_numpy(w.reshape(3, 3, 3, cout, cin)).permute(3, 4, 0, 1, 2).contiguous()
    xp = torch.nn.functional.pad(xt, (1, 1, 1, 1, 1, 1), mode="circular")
    yt = torch.nn.functional.conv3d(xp, wt, torch.from_numpy(bias)).permute(0, 2, 3, 4, 1).numpy()
    assert np.allclose(y, yt, rtol=1e-12, atol=1e-12)
