"""Replica ensembles: many independent periodic binary-fluid lattices of
uneven sizes advanced together by libludwig_hip.so (ldw_ensemble_*).

Every replica follows the arithmetic of a single lattice driven by
``ldw_step_binary`` (one :class:`ludwig_amd.BinaryFluid`) with the same
size, parameters and seed; one persistent launch per time step, replayed
from a HIP graph, advances all of them.  :func:`draw_replicas` produces the
uneven-size, random-parameter ensembles used as surrogate training data.
"""

import ctypes
import math

import numpy as np

from . import FORCE_STRESS_DIVERGENCE, _ck, _ptr, lib


class ReplicaSpec(ctypes.Structure):
    """ldw_replica_t (include/ludwig_hip.h)."""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("seed", ctypes.c_int),
                ("phi0", ctypes.c_double), ("noise", ctypes.c_double),
                ("a", ctypes.c_double), ("b", ctypes.c_double), ("kappa", ctypes.c_double),
                ("mobility", ctypes.c_double),
                ("eta_shear", ctypes.c_double), ("eta_bulk", ctypes.c_double),
                ("rho0", ctypes.c_double)]


def replica(shape, seed, phi0=0.0, noise=0.1, a=-0.0625, b=0.0625, kappa=0.04, mobility=0.25,
            tau=1.0, rho0=1.0):
    """One replica: lattice `shape` (int or 3-tuple), spinodal start
    phi0 + noise (U - 1/2) from `seed`, relaxation time tau (eta = (tau - 1/2)/3
    for shear and bulk)."""
    if np.isscalar(shape):
        shape = (int(shape),) * 3
    eta = (tau - 0.5) / 3.0
    return ReplicaSpec(int(shape[0]), int(shape[1]), int(shape[2]), int(seed), float(phi0),
                       float(noise), float(a), float(b), float(kappa), float(mobility), eta, eta,
                       float(rho0))


def max_stable_mobility(a, b, kappa, phimax=1.1):
    """Largest mobility for which the explicit Cahn-Hilliard step damps every
    lattice mode: the linearised update multiplies mode k by
    1 - M k2 (a + 3 b phi^2 + kappa k2), k2 = sum_a 2 (1 - cos k_a) <= 12,
    which must stay >= -1 for |phi| <= phimax."""
    return 2.0 / (12.0 * (a + 3.0 * b * phimax * phimax + 12.0 * kappa))


def draw_replicas(nrep, seed, sizes=(48, 64, 80, 96), phi0=(-0.3, 0.3), kappa=(0.02, 0.08),
                  mobility=(0.1, 0.5), noise=0.1, a=-0.0625, b=0.0625, tau=1.0,
                  stability_margin=0.95):
    """nrep replicas with cubic sizes drawn uniformly from `sizes`, phi0, kappa
    and M uniform in their ranges, and seeds from one generator.  A drawn
    mobility above `stability_margin` times the explicit scheme's limit for
    that kappa is clipped to it (the upper part of M in [0.1, 0.5] is unstable
    for the larger kappa)."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(nrep):
        n = int(rng.choice(sizes))
        p0 = float(rng.uniform(*phi0))
        k = float(rng.uniform(*kappa))
        m = float(rng.uniform(*mobility))
        m = min(m, stability_margin * max_stable_mobility(a, b, k))
        s = int(rng.integers(2, 2**31 - 1))
        out.append(replica(n, s, phi0=p0, noise=noise, a=a, b=b, kappa=k, mobility=m, tau=tau))
    return out


_SIG_DONE = False


def _bind():
    global _SIG_DONE
    L = lib()
    if _SIG_DONE:
        return L
    vp = ctypes.c_void_p
    sig = {
        "ldw_ensemble_create": [ctypes.c_int, ctypes.POINTER(ReplicaSpec), ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)],
        "ldw_ensemble_free": [vp],
        "ldw_ensemble_tune": [vp, ctypes.c_int],
        "ldw_ensemble_tune_results": [vp, ctypes.c_int, vp],
        "ldw_ensemble_config": [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)],
        "ldw_ensemble_config_set": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
        "ldw_ensemble_graph_set": [vp, ctypes.c_int],
        "ldw_ensemble_step": [vp, ctypes.c_int],
        "ldw_ensemble_snapshot": [vp, ctypes.c_int, vp, vp, vp],
        "ldw_ensemble_f_read": [vp, ctypes.c_int, vp],
        "ldw_ensemble_snapshot_device": [vp, ctypes.c_int, vp],
        "ldw_ensemble_conserved_sums": [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
        "ldw_ensemble_load_device": [vp, ctypes.c_int, vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.ldw_ensemble_time.argtypes = [vp]
    L.ldw_ensemble_time.restype = ctypes.c_longlong
    _SIG_DONE = True
    return L


class Ensemble:
    """A replica ensemble on the GPU (precision 32: FP32-shifted storage)."""

    def __init__(self, specs, precision=32, order=1, force_method=FORCE_STRESS_DIVERGENCE,
                 fatal=False):
        L = _bind()
        L.ldw_error_mode_set(1 if fatal else 0)
        self.specs = list(specs)
        self.shapes = [(s.nx, s.ny, s.nz) for s in self.specs]
        self.precision = precision
        arr = (ReplicaSpec * len(self.specs))(*self.specs)
        self.h = ctypes.c_void_p()
        _ck(L.ldw_ensemble_create(len(self.specs), arr, precision, order, force_method,
                                  ctypes.byref(self.h)), "ldw_ensemble_create")

    @property
    def nsite(self):
        return sum(int(np.prod(s)) for s in self.shapes)

    def tune(self, reps=3):
        L = _bind()
        _ck(L.ldw_ensemble_tune(self.h, reps), "ldw_ensemble_tune")
        n = L.ldw_ensemble_tune_results(self.h, 0, None)
        rows = np.zeros(5 * max(n, 1))
        L.ldw_ensemble_tune_results(self.h, n, rows.ctypes.data_as(ctypes.c_void_p))
        return rows[:5 * n].reshape(n, 5)

    def config(self):
        cfg = (ctypes.c_int * 4)()
        ms = ctypes.c_double()
        _ck(_bind().ldw_ensemble_config(self.h, cfg, ctypes.byref(ms)), "ldw_ensemble_config")
        return {"tile": (cfg[0], cfg[1]), "xplanes": cfg[2], "blocks": cfg[3], "tuned_ms": ms.value}

    def set_config(self, tz, ty, xplanes, blocks=0):
        _ck(_bind().ldw_ensemble_config_set(self.h, tz, ty, xplanes, blocks), "ldw_ensemble_config_set")

    def use_graph(self, on=True):
        _ck(_bind().ldw_ensemble_graph_set(self.h, 1 if on else 0), "ldw_ensemble_graph_set")

    def step(self, nsteps=1):
        _ck(_bind().ldw_ensemble_step(self.h, nsteps), "ldw_ensemble_step")

    @property
    def time(self):
        return _bind().ldw_ensemble_time(self.h)

    def snapshot(self, r):
        """(phi, rho, u) of replica r as FP64 arrays in the device layout."""
        n = int(np.prod(self.shapes[r]))
        phi, rho, u = np.empty(n), np.empty(n), np.empty(3 * n)
        _ck(_bind().ldw_ensemble_snapshot(self.h, r, _ptr(phi), _ptr(rho), _ptr(u)),
            "ldw_ensemble_snapshot")
        return phi, rho, u

    def f(self, r):
        n = int(np.prod(self.shapes[r]))
        f = np.empty(19 * n)
        _ck(_bind().ldw_ensemble_f_read(self.h, r, _ptr(f)), "ldw_ensemble_f_read")
        return f

    def snapshot_device(self, r, dst_ptr):
        """phi, ux, uy, uz of replica r as 4 * nsite floats at device address dst_ptr."""
        _ck(_bind().ldw_ensemble_snapshot_device(self.h, r, ctypes.c_void_p(dst_ptr)),
            "ldw_ensemble_snapshot_device")

    def load_device(self, r, src_ptr):
        """Replace replica r's phi, u by 4 * nsite floats at device address
        src_ptr; distributions become the local equilibrium."""
        _ck(_bind().ldw_ensemble_load_device(self.h, r, ctypes.c_void_p(src_ptr)),
            "ldw_ensemble_load_device")

    def conserved_sums(self, r):
        out = (ctypes.c_double * 2)()
        _ck(_bind().ldw_ensemble_conserved_sums(self.h, r, out), "ldw_ensemble_conserved_sums")
        return out[0], out[1]

    def close(self):
        if self.h:
            _bind().ldw_ensemble_free(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
