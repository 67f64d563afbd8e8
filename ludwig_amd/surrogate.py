"""Learned surrogate for the binary-fluid step: a 4-level 3D convolutional
U-Net that advances (phi, u) by a fixed number of lattice Boltzmann steps.

Every operation of training and inference runs on the GPU through
libludwig_hip.so's own kernels: BF16 MFMA implicit-GEMM convolutions with
periodic padding (forward, backward data, backward weight), SiLU, 2x average
pooling, nearest upsampling, channel concatenation, the loss and Adam (FP32
master weights, BF16 copies for the MFMA passes).  This module only orders
the kernel calls and owns the device buffers.

Network (channels-last, periodic lattices whose sides are multiples of 8):

    input 16 ch: phi, ux, uy, uz, (phi0, kappa, M) scaled, 9 zero channels
    e0a 16->32, e0b 32->32            | level 0 (n)
    pool, e1a 32->48, e1b 48->48      | level 1 (n/2)
    pool, e2a 48->64, e2b 64->64      | level 2 (n/4)
    pool, b0 64->96, b1 96->96        | level 3 (n/8)
    up, concat(96 + 64), d2 160->64   | level 2
    up, concat(64 + 48), d1 112->48   | level 1
    up, concat(48 + 32), d0 80->32    | level 0
    out 32->16 (no activation): channels 0..3 = scaled change of phi, u

All convolutions are 3x3x3 with bias; SiLU after every one but `out`.
1.26 M parameters.
"""

import ctypes
import math

import numpy as np

from . import DeviceArray, _ck, lib

C = (32, 48, 64, 96)
CIN = 16
COUT = 16

LAYERS = [("e0a", CIN, C[0]), ("e0b", C[0], C[0]),
          ("e1a", C[0], C[1]), ("e1b", C[1], C[1]),
          ("e2a", C[1], C[2]), ("e2b", C[2], C[2]),
          ("b0", C[2], C[3]), ("b1", C[3], C[3]),
          ("d2", C[3] + C[2], C[2]), ("d1", C[2] + C[1], C[1]), ("d0", C[1] + C[0], C[0]),
          ("out", C[0], COUT)]
LEVEL = {"e0a": 0, "e0b": 0, "e1a": 1, "e1b": 1, "e2a": 2, "e2b": 2, "b0": 3, "b1": 3,
         "d2": 2, "d1": 1, "d0": 0, "out": 0}


def _v(d):
    return ctypes.c_void_p(d.ptr if isinstance(d, DeviceArray) else d)


def _off(d, nbytes):
    """Device address `nbytes` into buffer d."""
    return ctypes.c_void_p(d.ptr + int(nbytes))


def _weight_mask():
    """Over the flat parameter vector: True for convolution weights, False
    for biases (each layer is 27 cout cin weights, then cout biases)."""
    keep = []
    for name, ci, co in LAYERS:
        keep += [np.ones(27 * co * ci, dtype=bool), np.zeros(co, dtype=bool)]
    return np.concatenate(keep)


def write_inference(path, w, step, norm_dict):
    """The inference file of a flat FP32 parameter vector w (host code,
    no device needed): BF16 weights, FP32 biases, the normaliser."""
    keep = _weight_mask()
    w = np.asarray(w, dtype=np.float32)
    if w.size != keep.size:
        raise ValueError(f"{w.size} parameters, the network has {keep.size}")
    u = w[keep].view(np.uint32).astype(np.uint64)
    bits = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    np.savez_compressed(path, w_bf16=bits, bias=w[~keep], step=np.array(step),
                        **{"norm_" + k: np.array(v) for k, v in norm_dict.items()})


def inference_from_training_file(src, dst, **gains):
    """The inference file of a save() file (weights + optimiser state),
    with the normaliser stored in it and, if given, the jump gains
    (jump_gain_phi, jump_gain_u) replaced."""
    d = np.load(src)
    norm = {k[5:]: float(d[k]) for k in d.files if k.startswith("norm_")}
    norm.update({k: float(v) for k, v in gains.items()})
    write_inference(dst, d["w"], int(d["step"]), norm)


class UNet:
    """Parameters, optimiser state and the forward/backward passes."""

    def __init__(self, seed=0, out_scale=0.1):
        self.layout = {}
        off = 0
        for name, ci, co in LAYERS:
            self.layout[name] = (off, ci, co)
            off += 27 * co * ci + co
        self.nparam = off
        rng = np.random.default_rng(seed)
        w = np.zeros(off, dtype=np.float32)
        for name, ci, co in LAYERS:
            o = self.layout[name][0]
            std = math.sqrt(2.0 / (27 * ci))
            if name == "out":
                std *= out_scale
            w[o:o + 27 * co * ci] = rng.standard_normal(27 * co * ci) * std
            if name == "e0a":
                # the 9 padding input channels never carry signal
                wl = w[o:o + 27 * co * ci].reshape(27, co, ci)
                wl[:, :, 7:] = 0.0
        self.w = DeviceArray.from_host(w, np.float32)
        self.g = DeviceArray(off, np.float32)
        self.m = DeviceArray.from_host(np.zeros(off, np.float32), np.float32)
        self.v = DeviceArray.from_host(np.zeros(off, np.float32), np.float32)
        self.wf, self.wd = {}, {}
        for name, ci, co in LAYERS:
            self.wf[name] = DeviceArray(27 * co * ci, np.uint16)
            self.wd[name] = DeviceArray(27 * co * ci, np.uint16)
        self.step_count = 0
        self._bufs = {}
        self.prepare()

    # -- parameters -------------------------------------------------------
    def wptr(self, name, arr=None):
        o = self.layout[name][0]
        return _off(arr or self.w, 4 * o)

    def bptr(self, name, arr=None):
        o, ci, co = self.layout[name]
        return _off(arr or self.w, 4 * (o + 27 * co * ci))

    def prepare(self):
        """BF16 weight copies for the forward and backward-data passes."""
        L = lib()
        for name, ci, co in LAYERS:
            _ck(L.ldw_conv3d_weights_prep(self.wptr(name), _v(self.wf[name]), _v(self.wd[name]), ci, co),
                "ldw_conv3d_weights_prep")

    def params(self):
        return self.w.download()

    def set_params(self, w):
        self.w.upload(np.asarray(w, dtype=np.float32))
        self.prepare()

    def grads(self):
        return self.g.download()

    def save(self, path, norm=None):
        """Weights and optimiser state (numpy .npz), with the normaliser the
        weights were trained under when it is given."""
        extra = {} if norm is None else {"norm_" + k: np.array(v) for k, v in norm.to_dict().items()}
        np.savez(path, w=self.w.download(), m=self.m.download(), v=self.v.download(),
                 step=np.array(self.step_count), **extra)

    def save_inference(self, path, norm):
        """What inference needs, compactly: the convolution weights as BF16
        (round to nearest even, the rounding the MFMA copies get anyway),
        biases in FP32, and the normaliser.  Loading it reproduces the
        network's outputs exactly."""
        write_inference(path, self.w.download(), self.step_count, norm.to_dict())

    def load(self, path):
        """Weights from save() or save_inference(); returns the normaliser
        stored with them, or None.  An inference file resets the optimiser
        state."""
        d = np.load(path)
        if "w_bf16" in d:
            keep = np.ones(self.nparam, dtype=bool)
            for name, ci, co in LAYERS:
                o = self.layout[name][0]
                keep[o + 27 * co * ci:o + 27 * co * ci + co] = False
            if d["w_bf16"].size != int(keep.sum()) or d["bias"].size != int((~keep).sum()):
                raise ValueError(f"{path}: parameter counts do not match the network")
            w = np.zeros(self.nparam, dtype=np.float32)
            w[keep] = (d["w_bf16"].astype(np.uint32) << 16).view(np.float32)
            w[~keep] = d["bias"]
            self.w.upload(w)
            z = np.zeros(self.nparam, dtype=np.float32)
            self.m.upload(z); self.v.upload(z)
        else:
            if d["w"].size != self.nparam:
                raise ValueError(f"{path}: {d['w'].size} parameters, the network has {self.nparam}")
            self.w.upload(d["w"]); self.m.upload(d["m"]); self.v.upload(d["v"])
        self.step_count = int(d["step"])
        self.prepare()
        if "norm_u_scale" in d:
            return Normaliser(**{k[5:]: float(d[k]) for k in d.files if k.startswith("norm_")})
        return None

    def grad_norm(self):
        out = ctypes.c_double()
        _ck(lib().ldw_nn_sumsq(_v(self.g), self.nparam, ctypes.byref(out)), "ldw_nn_sumsq")
        return math.sqrt(out.value)

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0, clip=None):
        """One Adam step; with `clip`, the gradient is first scaled to a
        global norm of at most `clip`."""
        self.step_count += 1
        gscale = 1.0
        if clip is not None:
            gn = self.grad_norm()
            if not math.isfinite(gn):
                raise FloatingPointError("non-finite gradient")
            gscale = min(1.0, clip / max(gn, 1e-30))
        _ck(lib().ldw_nn_adam(_v(self.w), _v(self.g), _v(self.m), _v(self.v), self.nparam, lr, b1, b2,
                              eps, wd, gscale, self.step_count), "ldw_nn_adam")
        self.prepare()

    # -- buffers ----------------------------------------------------------
    def buffers(self, batch, shape):
        key = (batch,) + tuple(shape)
        if key in self._bufs:
            return self._bufs[key]
        nx, ny, nz = shape
        if nx % 8 or ny % 8 or nz % 8:
            raise ValueError(f"lattice sides must be multiples of 8, got {shape}")
        P = [batch * (nx >> l) * (ny >> l) * (nz >> l) for l in range(4)]
        b = {"P": P, "shape": [(nx >> l, ny >> l, nz >> l) for l in range(4)]}
        for name, ci, co in LAYERS:
            p = P[LEVEL[name]]
            b["z_" + name] = DeviceArray(p * co, np.float32)
            if name != "out":
                b["a_" + name] = DeviceArray(p * co, np.uint16)
                b["da_" + name] = DeviceArray(p * co, np.float32)
            b["dz_" + name] = DeviceArray(p * co, np.float32)
            b["dzb_" + name] = DeviceArray(p * co, np.uint16)
        for l, c in ((1, C[0]), (2, C[1]), (3, C[2])):
            b[f"pool{l}"] = DeviceArray(P[l] * c, np.uint16)
            b[f"dpool{l}"] = DeviceArray(P[l] * c, np.float32)
        for l, c_up, c_skip in ((2, C[3], C[2]), (1, C[2], C[1]), (0, C[1], C[0])):
            b[f"dup{l}"] = DeviceArray(P[l] * c_up, np.float32)
            b[f"cat{l}"] = DeviceArray(P[l] * (c_up + c_skip), np.uint16)
            b[f"dcat{l}"] = DeviceArray(P[l] * (c_up + c_skip), np.float32)
        self._bufs[key] = b
        return b

    # -- passes -----------------------------------------------------------
    def _conv(self, b, name, x, act=True):
        """Convolution with SiLU fused into its epilogue; the pre-activation
        is kept only when a backward pass will follow."""
        L = lib()
        _, ci, co = self.layout[name]
        l = LEVEL[name]
        nx, ny, nz = b["shape"][l]
        batch = b["batch"]
        if act:
            z = _v(b["z_" + name]) if b["keep"] else None
            _ck(L.ldw_conv3d_fwd_silu(_v(x), _v(self.wf[name]), self.bptr(name), z, _v(b["a_" + name]),
                                      batch, nx, ny, nz, ci, co), f"conv {name}")
        else:
            _ck(L.ldw_conv3d_fwd(_v(x), _v(self.wf[name]), self.bptr(name), _v(b["z_" + name]), batch,
                                 nx, ny, nz, ci, co, 0), f"conv {name}")

    def forward(self, x, batch, shape, keep=True):
        """x: device BF16 [batch*nsite][16].  Returns the FP32 output buffer
        [batch*nsite][16] (channels 0..3 used).  keep=False (inference)
        skips storing what only the backward pass needs."""
        L = lib()
        b = self.buffers(batch, shape)
        b["keep"] = keep
        b["batch"] = batch
        b["x"] = x
        sh = b["shape"]
        self._conv(b, "e0a", x)
        self._conv(b, "e0b", b["a_e0a"])
        for l, prev, na, nb in ((1, "e0b", "e1a", "e1b"), (2, "e1b", "e2a", "e2b"), (3, "e2b", "b0", "b1")):
            c = self.layout[prev][2]
            _ck(L.ldw_nn_pool2_fwd(_v(b["a_" + prev]), _v(b[f"pool{l}"]), batch, *sh[l - 1], c), "pool")
            self._conv(b, na, b[f"pool{l}"])
            self._conv(b, nb, b["a_" + na])
        for l, low, skip, name in ((2, "b1", "e2b", "d2"), (1, "d2", "e1b", "d1"), (0, "d1", "e0b", "d0")):
            cu, cs = self.layout[low][2], self.layout[skip][2]
            _ck(L.ldw_nn_up2_concat(_v(b["a_" + low]), cu, _v(b["a_" + skip]), cs, _v(b[f"cat{l}"]), batch,
                                    *sh[l]), "up + concat")
            self._conv(b, name, b[f"cat{l}"])
        self._conv(b, "out", b["a_d0"], act=False)
        self._last = b
        return b["z_out"]

    def _conv_back(self, b, name, x, dzb, dz, dx=None, acc_dx=0):
        """Weight and bias gradients of layer `name`; input gradient into dx."""
        L = lib()
        _, ci, co = self.layout[name]
        l = LEVEL[name]
        nx, ny, nz = b["shape"][l]
        batch = b["batch"]
        _ck(L.ldw_conv3d_bwd_weight(_v(x), _v(dzb), self.wptr(name, self.g), batch, nx, ny, nz, ci, co, 0),
            f"wgrad {name}")
        _ck(L.ldw_nn_bias_grad(_v(dz), self.bptr(name, self.g), b["P"][l], co, 0), f"bias grad {name}")
        if dx is not None:
            _ck(L.ldw_conv3d_bwd_data(_v(dzb), _v(self.wd[name]), _v(dx), batch, nx, ny, nz, ci, co, acc_dx),
                f"dgrad {name}")

    def _act_back(self, b, name):
        L = lib()
        co = self.layout[name][2]
        _ck(L.ldw_nn_silu_bwd(_v(b["z_" + name]), _v(b["da_" + name]), _v(b["dz_" + name]),
                              _v(b["dzb_" + name]), b["P"][LEVEL[name]] * co), f"silu back {name}")

    def backward(self, dyb, dyf):
        """Gradients of the loss into self.g, from the output gradient dyb
        (BF16) and dyf (FP32), both [batch*nsite][16]."""
        L = lib()
        b = self._last
        if not b["keep"]:
            raise RuntimeError("backward after an inference-only forward pass")
        batch, sh, P = b["batch"], b["shape"], b["P"]
        self._conv_back(b, "out", b["a_d0"], dyb, dyf, b["da_d0"])
        for l, low, skip, name in ((0, "d1", "e0b", "d0"), (1, "d2", "e1b", "d1"), (2, "b1", "e2b", "d2")):
            cu, cs = self.layout[low][2], self.layout[skip][2]
            self._act_back(b, name)
            self._conv_back(b, name, b[f"cat{l}"], b["dzb_" + name], b["dz_" + name], b[f"dcat{l}"])
            # the skip branch's gradient is written first; pooling adds to it later
            _ck(L.ldw_nn_split(_v(b[f"dcat{l}"]), cu, cs, _v(b[f"dup{l}"]), _v(b["da_" + skip]), P[l], 0, 0),
                "split")
            _ck(L.ldw_nn_up2_bwd(_v(b[f"dup{l}"]), _v(b["da_" + low]), batch, *sh[l + 1], cu), "up back")
        for l, prev, na, nb in ((3, "e2b", "b0", "b1"), (2, "e1b", "e2a", "e2b"), (1, "e0b", "e1a", "e1b")):
            self._act_back(b, nb)
            self._conv_back(b, nb, b["a_" + na], b["dzb_" + nb], b["dz_" + nb], b["da_" + na])
            self._act_back(b, na)
            self._conv_back(b, na, b[f"pool{l}"], b["dzb_" + na], b["dz_" + na], b[f"dpool{l}"])
            c = self.layout[prev][2]
            _ck(L.ldw_nn_pool2_bwd(_v(b[f"dpool{l}"]), _v(b["da_" + prev]), batch, *sh[l - 1], c, 1),
                "pool back")
        self._act_back(b, "e0b")
        self._conv_back(b, "e0b", b["a_e0a"], b["dzb_e0b"], b["dz_e0b"], b["da_e0a"])
        self._act_back(b, "e0a")
        self._conv_back(b, "e0a", b["x"], b["dzb_e0a"], b["dz_e0a"], None)


class Normaliser:
    """Scales of the network's inputs and outputs, from training data.

    jump_gain_phi, jump_gain_u: factors on the predicted changes of phi
    and u when a jump is applied (decode), 1 by default.  A network trained
    on the mean-square error predicts the conditional mean of the change,
    which shrinks the part of a domain wall's displacement it cannot be
    sure of, so its jumps advance the domain structure a little less than
    the 250 steps they stand for; the gain is a scalar calibrated on
    training replicas (surrogate_study.calibrate), never on held-out ones."""

    def __init__(self, u_scale, dphi_scale, du_scale, phi_shift=0.0, phi_scale=1.0, jump_gain_phi=1.0,
                 jump_gain_u=1.0):
        self.u_scale, self.dphi_scale, self.du_scale = float(u_scale), float(dphi_scale), float(du_scale)
        self.phi_shift, self.phi_scale = float(phi_shift), float(phi_scale)
        self.jump_gain_phi, self.jump_gain_u = float(jump_gain_phi), float(jump_gain_u)

    @staticmethod
    def params(spec):
        """The three scalar input channels of a replica (phi0, kappa, M)."""
        return (spec.phi0 / 0.3, (spec.kappa - 0.05) / 0.03, (spec.mobility - 0.3) / 0.2)

    def norm(self, spec):
        return (ctypes.c_float * 6)(self.phi_shift, self.phi_scale, self.u_scale, *self.params(spec))

    def to_dict(self):
        return dict(u_scale=self.u_scale, dphi_scale=self.dphi_scale, du_scale=self.du_scale,
                    phi_shift=self.phi_shift, phi_scale=self.phi_scale, jump_gain_phi=self.jump_gain_phi,
                    jump_gain_u=self.jump_gain_u)

    def with_gains(self, phi=1.0, u=1.0):
        return Normaliser(self.u_scale, self.dphi_scale, self.du_scale, self.phi_shift, self.phi_scale, phi, u)


def encode(state, nsite, spec, normaliser, out, offset_sites=0):
    """Network input of one sample (4 SoA planes of floats at `state`) into
    the BF16 batch buffer `out`, at sample offset `offset_sites`."""
    _ck(lib().ldw_nn_encode(_v(state), nsite, normaliser.norm(spec), _off(out, 2 * CIN * offset_sites), CIN),
        "ldw_nn_encode")


def target(now, nxt, nsite, normaliser, out, offset_sites=0):
    _ck(lib().ldw_nn_target(_v(now), _v(nxt), nsite, _off(out, 4 * 4 * offset_sites),
                            normaliser.dphi_scale, normaliser.du_scale), "ldw_nn_target")


PHI_MAX, U_MAX = 1.5, 0.1      # clamps on a jump's result (|phi| ~ 1, |u| << 0.1)


def decode(now, y, nsite, normaliser, out, offset_sites=0):
    """The state after a jump: now + predicted change (times the
    normaliser's jump gains), clamped."""
    _ck(lib().ldw_nn_decode(_v(now), _off(y, 4 * COUT * offset_sites), COUT, nsite, _v(out),
                            normaliser.dphi_scale / normaliser.jump_gain_phi,
                            normaliser.du_scale / normaliser.jump_gain_u, PHI_MAX, U_MAX), "ldw_nn_decode")


# The 48 elements of the cubic group (axis permutations x reflections).
# The lattice dynamics commute with each of them (the D3Q19 velocity set,
# the centred stencils and the upwind fluxes are all invariant), so a
# training pair mapped by one is again a pair of the dynamics.
CUBIC_GROUP = [(p, (f >> 2 & 1, f >> 1 & 1, f & 1))
               for p in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)) for f in range(8)]


def cubic(src, dst, n, perm, flip):
    """dst = the state src (4 planes of n^3 floats) under one cubic map."""
    _ck(lib().ldw_nn_cubic(_v(src), _v(dst), n, (ctypes.c_int * 3)(*perm), (ctypes.c_int * 3)(*flip)),
        "ldw_nn_cubic")


def loss_and_grad(y, tgt, npos, dyb, dyf, weights=(1.0, 1.0, 1.0, 1.0)):
    loss = ctypes.c_double()
    w = (ctypes.c_float * 4)(*weights)
    _ck(lib().ldw_nn_mse(_v(y), COUT, _v(tgt), npos, _v(dyb), _v(dyf), w, ctypes.byref(loss)), "ldw_nn_mse")
    return loss.value


# ---------------------------------------------------------------------------
# Data, training and jumps


class Dataset:
    """(phi, u) snapshots of every replica of an ensemble, kept on the
    device: snaps[r][k] is replica r at step k*every, 4 SoA planes of floats.
    With offset c > 0 also osnaps[r][k], replica r at step k*every + c: the
    times a jump pipeline with c correction steps reaches."""

    def __init__(self, ensemble, nsnap, every, replicas=None, offset=0):
        self.specs = ensemble.specs
        self.shapes = ensemble.shapes
        self.every = every
        self.offset = int(offset)
        if not 0 <= self.offset < every:
            raise ValueError(f"offset {offset} outside [0, {every})")
        self.replicas = list(range(len(self.specs))) if replicas is None else list(replicas)
        self.snaps = {r: [] for r in self.replicas}
        self.osnaps = {r: [] for r in self.replicas} if self.offset else None
        self._take(ensemble, self.snaps)
        for k in range(nsnap):
            if self.offset:
                ensemble.step(self.offset)
                self._take(ensemble, self.osnaps)
                ensemble.step(every - self.offset)
            else:
                ensemble.step(every)
            self._take(ensemble, self.snaps)

    def _take(self, e, into):
        for r in self.replicas:
            d = DeviceArray(4 * int(np.prod(self.shapes[r])), np.float32)
            e.snapshot_device(r, d.ptr)
            into[r].append(d)

    def pairs(self, lag=1, kmin=0):
        """(replica, snapshot) of every pair k -> k + lag with k >= kmin."""
        return [(r, k) for r in self.replicas for k in range(kmin, len(self.snaps[r]) - lag)]

    def normaliser(self, sample=16, lag=1, seed=0, kmin=0):
        """Scales from a sample of pairs: u to unit RMS, and the changes of
        phi and u over one jump to unit RMS (medians over the sample)."""
        rng = np.random.default_rng(seed)
        pr = self.pairs(lag, kmin)
        pick = [pr[i] for i in rng.choice(len(pr), min(sample, len(pr)), replace=False)]
        us, dps, dus = [], [], []
        for r, k in pick:
            a = self.snaps[r][k].download().reshape(4, -1)
            b = self.snaps[r][k + lag].download().reshape(4, -1)
            us.append(np.sqrt(np.mean(a[1:] ** 2)))
            dps.append(np.sqrt(np.mean((b[0] - a[0]) ** 2)))
            dus.append(np.sqrt(np.mean((b[1:] - a[1:]) ** 2)))
        rms = lambda v: max(float(np.median(v)), 1e-12)
        return Normaliser(1.0 / rms(us), 1.0 / rms(dps), 1.0 / rms(dus))

    def free(self):
        for series in (self.snaps, self.osnaps or {}):
            for r in series:
                for d in series[r]:
                    d.free()
        self.snaps, self.osnaps = {}, None


class Trainer:
    """Adam on the GPU over batches of same-size pairs (snapshot k -> k + lag)."""

    def __init__(self, net, normaliser, batch=4, weights=(1.0, 0.1, 0.1, 0.1), clip=1.0, augment=False,
                 seed=0):
        """weights: loss weights of (phi, ux, uy, uz) changes -- the velocity
        over a whole jump is nearly decorrelated, so phi leads; clip: global
        gradient-norm bound; augment: map every training pair of a cubic
        lattice by a random element of the cubic group (CUBIC_GROUP)."""
        self.net, self.norm, self.batch = net, normaliser, batch
        self.weights, self.clip = tuple(weights), clip
        self.augment = augment
        self._rng = np.random.default_rng(seed + 7919)
        self._io = {}
        self._aug = {}

    def _bufs(self, shape, batch):
        key = (batch,) + tuple(shape)
        if key not in self._io:
            n = batch * int(np.prod(shape))
            self._io[key] = dict(x=DeviceArray(n * CIN, np.uint16), t=DeviceArray(n * 4, np.float32),
                                 dyb=DeviceArray(n * COUT, np.uint16), dyf=DeviceArray(n * COUT, np.float32))
        return self._io[key]

    def groups(self, data, pairs):
        by = {}
        for r, k in pairs:
            by.setdefault(tuple(data.shapes[r]), []).append((r, k))
        return by

    def step(self, data, items, lr, lag=1, train=True, zero=False):
        """One optimiser step (or, train=False, only the loss) on `items`,
        a list of (replica, snapshot) of one lattice size.  zero=True scores
        the zero-change prediction instead of the network."""
        return self.step_arrays([(data.snaps[r][k], data.snaps[r][k + lag], data.specs[r])
                                 for r, k in items], tuple(data.shapes[items[0][0]]), lr, train, zero)

    def step_arrays(self, items, shape, lr, train=True, zero=False):
        """The same on explicit (state now, state one jump later, replica
        spec) device arrays of one lattice shape."""
        ns = int(np.prod(shape))
        io = self._bufs(shape, len(items))
        if train and self.augment and shape[0] == shape[1] == shape[2]:
            key = (len(items),) + tuple(shape)
            if key not in self._aug:
                self._aug[key] = [(DeviceArray(4 * ns, np.float32), DeviceArray(4 * ns, np.float32))
                                  for _ in range(len(items))]
            mapped = []
            for j, (now, nxt, spec) in enumerate(items):
                perm, flip = CUBIC_GROUP[self._rng.integers(len(CUBIC_GROUP))]
                a, b = self._aug[key][j]
                cubic(now, a, shape[0], perm, flip)
                cubic(nxt, b, shape[0], perm, flip)
                mapped.append((a, b, spec))
            items = mapped
        for j, (now, nxt, spec) in enumerate(items):
            encode(now, ns, spec, self.norm, io["x"], j * ns)
            target(now, nxt, ns, self.norm, io["t"], j * ns)
        if zero:
            if "zero" not in io:
                io["zero"] = DeviceArray.from_host(np.zeros(len(items) * ns * COUT, np.float32), np.float32)
            return loss_and_grad(io["zero"], io["t"], len(items) * ns, io["dyb"], io["dyf"], self.weights)
        y = self.net.forward(io["x"], len(items), shape)
        loss = loss_and_grad(y, io["t"], len(items) * ns, io["dyb"], io["dyf"], self.weights)
        if train:
            self.net.backward(io["dyb"], io["dyf"])
            self.net.adam(lr, clip=self.clip)
        return loss

    def fit(self, data, pairs, iters, lr=1e-3, lag=1, seed=0, lr_final=None, warmup=50, log=None,
            max_seconds=None):
        """`iters` Adam steps (fewer if `max_seconds` of wall time run out;
        the loss read-back each step keeps the host in step with the GPU)."""
        import time
        t_start = time.perf_counter()
        rng = np.random.default_rng(seed)
        groups = [g for g in self.groups(data, pairs).values()]
        weights = np.array([len(g) for g in groups], dtype=float)
        weights /= weights.sum()
        losses = []
        for it in range(iters):
            g = groups[rng.choice(len(groups), p=weights)]
            items = [g[i] for i in rng.choice(len(g), min(self.batch, len(g)), replace=False)]
            # exponential decay over the iterations or, with a time budget,
            # over the budget (so a run stopped by time ends at lr_final)
            frac = it / max(iters - 1, 1)
            if max_seconds is not None:
                frac = max(frac, (time.perf_counter() - t_start) / max_seconds)
            lr_it = lr if lr_final is None else lr * (lr_final / lr) ** min(frac, 1.0)
            lr_it *= min(1.0, (it + 1) / max(warmup, 1))
            losses.append(self.step(data, items, lr_it, lag))
            if log and (it % 100 == 0 or it == iters - 1):
                log(f"iter {it} loss {losses[-1]:.4e} ({time.perf_counter() - t_start:.1f} s)")
            if max_seconds is not None and time.perf_counter() - t_start > max_seconds:
                if log:
                    log(f"stopped after {it + 1} iterations: time budget {max_seconds} s")
                break
        return losses

    def evaluate(self, data, pairs, lag=1, zero=False):
        return float(np.mean([self.step(data, [p], 0.0, lag, train=False, zero=zero) for p in pairs]))

    def fit_pushforward(self, data, pf, pairs, iters, lr, lr_final=None, p_pf=0.5, refresh_every=400,
                        seed=1, log=None, max_seconds=None):
        """Adam steps on a mix of the plain pairs (snapshot k -> k + 1) and,
        with probability p_pf, pushforward pairs: the state the jump
        pipeline reaches from snapshot k (one network jump, then the
        correction steps, pf.inputs[r][k]) -> the true state one jump later
        (data.osnaps[r][k + 2]).  The pushforward inputs are rebuilt with
        the current network every `refresh_every` iterations."""
        import time
        t_start = time.perf_counter()
        rng = np.random.default_rng(seed)
        plain = [g for g in self.groups(data, pairs).values()]
        pfp = [g for g in self.groups(data, pf.pairs()).values()]
        wp = np.array([len(g) for g in plain], float); wp /= wp.sum()
        wf = np.array([len(g) for g in pfp], float); wf /= wf.sum()
        losses, losses_pf = [], []
        for it in range(iters):
            if it % refresh_every == 0:
                t0 = time.perf_counter()
                pf.refresh(self.net, self.norm)
                if log:
                    log(f"iter {it}: pushforward inputs rebuilt in {time.perf_counter() - t0:.1f} s")
            frac = it / max(iters - 1, 1)
            if max_seconds is not None:
                frac = max(frac, (time.perf_counter() - t_start) / max_seconds)
            lr_it = lr if lr_final is None else lr * (lr_final / lr) ** min(frac, 1.0)
            if rng.random() < p_pf:
                g = pfp[rng.choice(len(pfp), p=wf)]
                sel = [g[i] for i in rng.choice(len(g), min(self.batch, len(g)), replace=False)]
                items = [(pf.inputs[r][k], data.osnaps[r][k + 2], data.specs[r]) for r, k in sel]
                losses_pf.append(self.step_arrays(items, tuple(data.shapes[sel[0][0]]), lr_it))
            else:
                g = plain[rng.choice(len(plain), p=wp)]
                sel = [g[i] for i in rng.choice(len(g), min(self.batch, len(g)), replace=False)]
                losses.append(self.step(data, sel, lr_it))
            if log and it % 100 == 0:
                lp = np.mean(losses[-50:]) if losses else float("nan")
                lf = np.mean(losses_pf[-50:]) if losses_pf else float("nan")
                log(f"iter {it} loss {lp:.4e} pushforward loss {lf:.4e} "
                    f"({time.perf_counter() - t_start:.1f} s)")
            if max_seconds is not None and time.perf_counter() - t_start > max_seconds:
                if log:
                    log(f"stopped after {it + 1} iterations: time budget {max_seconds} s")
                break
        return losses, losses_pf


class Pushforward:
    """The inputs the jump pipeline feeds the network after its first jump:
    for each replica r and snapshot k, one network jump from data.snaps[r][k]
    (to step (k + 1) every), then data.offset plain lattice Boltzmann
    correction steps, giving the state at step (k + 1) every + offset.  The
    true state one jump later is data.osnaps[r][k + 2].  The replicas are
    advanced by an ensemble of their own (FP32-shifted storage)."""

    def __init__(self, data, replicas, kmin, conserve=True, chunk=4):
        from .ensemble import Ensemble
        if not data.offset:
            raise ValueError("pushforward pairs need a dataset with offset snapshots")
        self.data = data
        self.replicas = list(replicas)
        self.kmin = kmin
        self.kmax = len(data.osnaps[self.replicas[0]]) - 3    # osnaps[k + 2] must exist
        if self.kmax < kmin:
            raise ValueError("too few snapshots for a pushforward pair")
        self.conserve, self.chunk = conserve, chunk
        self.ens = Ensemble([data.specs[r] for r in self.replicas], precision=32, fatal=True)
        self.ens.tune(reps=1)
        self.slot = {r: j for j, r in enumerate(self.replicas)}
        self.inputs = {r: {k: DeviceArray(4 * int(np.prod(data.shapes[r])), np.float32)
                           for k in range(kmin, self.kmax + 1)} for r in self.replicas}
        self._tmp = {}

    def pairs(self):
        return [(r, k) for r in self.replicas for k in range(self.kmin, self.kmax + 1)]

    def refresh(self, net, norm):
        L = lib()
        data = self.data
        by = {}
        for r in self.replicas:
            by.setdefault(tuple(data.shapes[r]), []).append(r)
        for k in range(self.kmin, self.kmax + 1):
            for shape, rs in by.items():
                ns = int(np.prod(shape))
                for c0 in range(0, len(rs), self.chunk):
                    part = rs[c0:c0 + self.chunk]
                    key = (len(part),) + shape
                    if key not in self._tmp:
                        self._tmp[key] = (DeviceArray(len(part) * ns * CIN, np.uint16),
                                          DeviceArray(4 * ns, np.float32))
                    x, out = self._tmp[key]
                    for j, r in enumerate(part):
                        encode(data.snaps[r][k], ns, data.specs[r], norm, x, j * ns)
                    y = net.forward(x, len(part), shape, keep=False)
                    for j, r in enumerate(part):
                        decode(data.snaps[r][k], y, ns, norm, out, j * ns)
                        if self.conserve:
                            _ck(L.ldw_nn_restore_sums(_v(data.snaps[r][k]), _v(out), ns, 4),
                                "ldw_nn_restore_sums")
                        self.ens.load_device(self.slot[r], out.ptr)
            self.ens.step(data.offset)
            for r in self.replicas:
                self.ens.snapshot_device(self.slot[r], self.inputs[r][k].ptr)

    def free(self):
        for r in self.inputs:
            for d in self.inputs[r].values():
                d.free()
        for x, out in self._tmp.values():
            x.free()
            out.free()
        self.inputs, self._tmp = {}, {}
        self.ens.close()


class Jumper:
    """Surrogate jumps of whole ensembles: every replica's (phi, u) is
    replaced by the network's prediction, distributions by the local
    equilibrium; plain lattice Boltzmann steps then correct it.

    u_mode: "pred" takes the predicted velocity; "rescale" rescales it to
    the root-mean-square speed before the jump (a mean-square-trained
    prediction shrinks the part of u it cannot predict); "keep" keeps the
    velocity from before the jump.

    conserve: shift the predicted phi and each velocity component by a
    constant so that their sums over the replica equal those before the
    jump.  The dynamics conserve sum phi exactly (flux form) and the total
    momentum (no net force), which is sum u to within the density
    variation; the network conserves neither."""

    def __init__(self, net, normaliser, u_mode="pred", conserve=False):
        if u_mode not in ("pred", "rescale", "keep"):
            raise ValueError(f"u_mode {u_mode!r}")
        self.net, self.norm, self.u_mode = net, normaliser, u_mode
        self.conserve = conserve
        self._io = {}

    def _bufs(self, shape, batch):
        key = (batch,) + tuple(shape)
        if key not in self._io:
            ns = int(np.prod(shape))
            self._io[key] = dict(x=DeviceArray(batch * ns * CIN, np.uint16),
                                 now=[DeviceArray(4 * ns, np.float32) for _ in range(batch)],
                                 out=DeviceArray(4 * ns, np.float32))
        return self._io[key]

    def jump(self, ensemble, replicas=None):
        """One network jump of the listed replicas (default: all), batched by size."""
        reps = list(range(len(ensemble.specs))) if replicas is None else list(replicas)
        by = {}
        for r in reps:
            by.setdefault(tuple(ensemble.shapes[r]), []).append(r)
        for shape, rs in by.items():
            ns = int(np.prod(shape))
            io = self._bufs(shape, len(rs))
            for j, r in enumerate(rs):
                ensemble.snapshot_device(r, io["now"][j].ptr)
                encode(io["now"][j], ns, ensemble.specs[r], self.norm, io["x"], j * ns)
            y = self.net.forward(io["x"], len(rs), shape, keep=False)
            for j, r in enumerate(rs):
                decode(io["now"][j], y, ns, self.norm, io["out"], j * ns)
                if self.u_mode == "keep":
                    _ck(lib().ldw_memcpy(_off(io["out"], 4 * ns), _off(io["now"][j], 4 * ns), 12 * ns, 3),
                        "ldw_memcpy")
                elif self.u_mode == "rescale":
                    a, b = ctypes.c_double(), ctypes.c_double()
                    _ck(lib().ldw_nn_sumsq(_off(io["now"][j], 4 * ns), 3 * ns, ctypes.byref(a)), "sumsq")
                    _ck(lib().ldw_nn_sumsq(_off(io["out"], 4 * ns), 3 * ns, ctypes.byref(b)), "sumsq")
                    if b.value > 0.0:
                        _ck(lib().ldw_nn_scale(_off(io["out"], 4 * ns), 3 * ns, math.sqrt(a.value / b.value)),
                            "ldw_nn_scale")
                if self.conserve:
                    _ck(lib().ldw_nn_restore_sums(_v(io["now"][j]), _v(io["out"]), ns, 4),
                        "ldw_nn_restore_sums")
                ensemble.load_device(r, io["out"].ptr)
