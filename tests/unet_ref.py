"""FP64 CPU reference of the surrogate U-Net (ludwig_amd/surrogate.py) for
the gradient checks: the same graph built from the FP64 convolution
checker (oracle.conv3_*) and numpy.  With quantise=True it rounds to BF16
where the GPU stores BF16 (weights, activations, pooled and output
gradients), so GPU and reference then differ only by FP32 accumulation."""

import numpy as np

import oracle as O
from bf16util import bf16_round
from ludwig_amd.surrogate import C, LAYERS, LEVEL


def silu(z):
    return z / (1 + np.exp(-z))


def dsilu(z):
    s = 1 / (1 + np.exp(-z))
    return s * (1 + z * (1 - s))


def pool(a, b, shape, c):
    nx, ny, nz = shape
    a5 = a.reshape(b, nx // 2, 2, ny // 2, 2, nz // 2, 2, c)
    return a5.mean(axis=(2, 4, 6)).ravel()


def pool_back(d, b, shape, c):
    nx, ny, nz = shape
    d4 = d.reshape(b, nx // 2, ny // 2, nz // 2, c) / 8
    return d4.repeat(2, 1).repeat(2, 2).repeat(2, 3).ravel()


def up(a, b, shape, c):            # shape: coarse sizes
    nx, ny, nz = shape
    return a.reshape(b, nx, ny, nz, c).repeat(2, 1).repeat(2, 2).repeat(2, 3).ravel()


def up_back(d, b, shape, c):       # shape: coarse sizes
    nx, ny, nz = shape
    return d.reshape(b, nx, 2, ny, 2, nz, 2, c).sum(axis=(2, 4, 6)).ravel()


def concat(a, ca, bb, cb, npos):
    return np.concatenate([a.reshape(npos, ca), bb.reshape(npos, cb)], axis=1).ravel()


class RefUNet:
    def __init__(self, params, quantise=True):
        self.q = (lambda v: bf16_round(v)) if quantise else (lambda v: v)
        # FP32 storage of pre-activations, biases and gradients (GPU side)
        self.f = (lambda v: np.asarray(v, np.float32).astype(np.float64)) if quantise else (lambda v: v)
        self.layout = {}
        off = 0
        for name, ci, co in LAYERS:
            self.layout[name] = (off, ci, co)
            off += 27 * co * ci + co
        self.p = np.asarray(params, dtype=np.float64)

    def wb(self, name):
        o, ci, co = self.layout[name]
        w = self.q(self.f(self.p[o:o + 27 * co * ci]))
        return w, self.f(self.p[o + 27 * co * ci:o + 27 * co * ci + co])

    def forward(self, x, batch, shape):
        q = self.q
        sh = [tuple(s >> l for s in shape) for l in range(4)]
        cache = {"x": x, "batch": batch, "sh": sh}

        def conv(name, inp, act=True):
            w, bias = self.wb(name)
            _, ci, co = self.layout[name]
            z = self.f(O.conv3_fwd(inp, w, bias, sh[LEVEL[name]], ci, co, batch))
            cache["in_" + name] = inp
            cache["z_" + name] = z
            if act:
                cache["a_" + name] = q(self.f(silu(z)))
                return cache["a_" + name]
            return z

        a = conv("e0b", conv("e0a", x))
        for l, prev, na, nb in ((1, "e0b", "e1a", "e1b"), (2, "e1b", "e2a", "e2b"), (3, "e2b", "b0", "b1")):
            c = self.layout[prev][2]
            p = q(self.f(pool(cache["a_" + prev], batch, sh[l - 1], c)))
            conv(nb, conv(na, p))
        low = "b1"
        for l, skip, name in ((2, "e2b", "d2"), (1, "e1b", "d1"), (0, "e0b", "d0")):
            cu, cs = self.layout[low][2], self.layout[skip][2]
            npos = batch * int(np.prod(sh[l]))
            u = up(cache["a_" + low], batch, sh[l + 1], cu)
            conv(name, concat(u, cu, cache["a_" + skip], cs, npos))
            low = name
        y = conv("out", cache["a_d0"], act=False)
        self.cache = cache
        return y

    def backward(self, dy):
        """dy: output gradient (FP64, as the GPU's FP32/BF16 pair rounds it)."""
        q, cache = self.q, self.cache
        f = self.f
        batch, sh = cache["batch"], cache["sh"]
        g = np.zeros_like(self.p)
        da = {}

        def conv_back(name, dz, want_dx=True):
            o, ci, co = self.layout[name]
            w, _ = self.wb(name)
            dzb = q(self.f(dz))
            l = LEVEL[name]
            npos = batch * int(np.prod(sh[l]))
            g[o:o + 27 * co * ci] = O.conv3_bwd_weight(cache["in_" + name], dzb, sh[l], ci, co, batch)
            g[o + 27 * co * ci:o + 27 * co * ci + co] = dz.reshape(npos, co).sum(0)
            if want_dx:
                return O.conv3_bwd_data(dzb, w, sh[l], ci, co, batch)
            return None

        def act_back(name):
            return self.f(da[name] * dsilu(cache["z_" + name]))

        da["d0"] = conv_back("out", dy)
        low_of = {"d0": "d1", "d1": "d2", "d2": "b1"}
        skip_of = {"d0": "e0b", "d1": "e1b", "d2": "e2b"}
        for l, name in ((0, "d0"), (1, "d1"), (2, "d2")):
            low, skip = low_of[name], skip_of[name]
            cu, cs = self.layout[low][2], self.layout[skip][2]
            npos = batch * int(np.prod(sh[l]))
            dcat = conv_back(name, act_back(name)).reshape(npos, cu + cs)
            da[skip] = dcat[:, cu:].ravel()
            da[low] = up_back(dcat[:, :cu].ravel(), batch, sh[l + 1], cu)
        for l, prev, na, nb in ((3, "e2b", "b0", "b1"), (2, "e1b", "e2a", "e2b"), (1, "e0b", "e1a", "e1b")):
            da[na] = conv_back(nb, act_back(nb))
            dp = conv_back(na, act_back(na))
            c = self.layout[prev][2]
            da[prev] = da[prev] + pool_back(dp, batch, sh[l - 1], c)
        da["e0a"] = conv_back("e0b", act_back("e0b"))
        conv_back("e0a", act_back("e0a"), want_dx=False)
        return g
