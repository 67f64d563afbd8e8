"""CPU oracle for the D3Q19 binary-fluid step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker (or the timed CPU
baseline).  The product library ``libludwig_hip.so`` never calls it.

It wraps ``liboracle.so`` (``ldw_oracle.c``), a C restatement of the
reference's per-site arithmetic in the reference's operation order.  Arrays
use the device layout: a periodic ``nx*ny*nz`` lattice, site
``(x*ny + y)*nz + z``, component-major multi-component fields.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

M10, BGK, TRT = 0, 1, 2

_lib = None


class LbParam(ctypes.Structure):
    _fields_ = [("relax", ctypes.c_int),
                ("eta_shear", ctypes.c_double),
                ("eta_bulk", ctypes.c_double),
                ("rho0", ctypes.c_double),
                ("fglobal", ctypes.c_double * 3)]


class FeParam(ctypes.Structure):
    _fields_ = [("a", ctypes.c_double), ("b", ctypes.c_double),
                ("kappa", ctypes.c_double), ("mobility", ctypes.c_double),
                ("order", ctypes.c_int), ("force_method", ctypes.c_int),
                ("stencil", ctypes.c_int), ("conserve", ctypes.c_int),
                ("grad_mu", ctypes.c_double * 3)]


class LeParam(ctypes.Structure):
    _fields_ = [("nplane", ctypes.c_int), ("nhalo", ctypes.c_int),
                ("uy", ctypes.c_double), ("nt0", ctypes.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("liboracle.so is not built: run make -C oracle oracle")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.ora_ns_uniform.restype = ctypes.c_uint32
        _lib.ora_kahan_sum.restype = ctypes.c_double
        _lib.ora_klein_sum.restype = ctypes.c_double
        _lib.ora_kahan_sum.argtypes = [ctypes.c_int64, ctypes.c_void_p]
        _lib.ora_phi_init_sum.restype = ctypes.c_double
        _lib.ora_phi_init_sum.argtypes = [ctypes.c_int64, ctypes.c_void_p]
        _lib.ora_klein_sum.argtypes = [ctypes.c_int64, ctypes.c_void_p]
        vp, d, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        _lib.ora_le_force.argtypes = [i, i, i, vp, vp, d, vp, vp]
        _lib.ora_le_cahn_hilliard.argtypes = [i, i, i, vp, vp, d, vp, vp, vp, d]
        _lib.ora_le_lb_bc.argtypes = [i, i, i, vp, d, vp]
        _lib.ora_le_shear_profile.argtypes = [i, i, i, vp, d, d, vp]
        _lib.ora_le_binary_steps.argtypes = [i, i, i, i, i, i, vp, vp, vp, vp, vp, vp, vp, vp, d]
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def lb_param(tau=None, eta=None, eta_bulk=None, relax=M10, rho0=1.0,
             fglobal=(0.0, 0.0, 0.0)):
    """Parameters as the reference derives them: ``viscosity`` sets shear
    and bulk (physics_rt.c:99-106); tau = 3 eta + 1/2."""
    if eta is None:
        eta = (tau - 0.5) / 3.0
    if eta_bulk is None:
        eta_bulk = eta
    p = LbParam()
    p.relax = relax
    p.eta_shear = eta
    p.eta_bulk = eta_bulk
    p.rho0 = rho0
    for i in range(3):
        p.fglobal[i] = fglobal[i]
    return p


def fe_param(a=-0.0625, b=0.0625, kappa=0.04, mobility=0.45, order=1,
             force_method=0, stencil=7, conserve=0, grad_mu=(0.0, 0.0, 0.0)):
    return FeParam(a, b, kappa, mobility, order, force_method, stencil, conserve,
                   (ctypes.c_double * 3)(*grad_mu))


def tables():
    cv = (ctypes.c_int * 57)()
    wv = np.zeros(19)
    ma = np.zeros((19, 19))
    mi = np.zeros((19, 19))
    lib().ora_d3q19_tables(cv, _p(wv), _p(ma), _p(mi))
    return np.array(list(cv), dtype=np.int64).reshape(19, 3), wv, ma, mi


def relaxation_rates(p):
    r = np.zeros(19)
    lib().ora_relaxation_rates(ctypes.byref(p), _p(r))
    return r


def collide(shape, p, f, F=None):
    nx, ny, nz = shape
    n = nx * ny * nz
    f = np.ascontiguousarray(f, dtype=np.float64).copy()
    rho = np.zeros(n)
    u = np.zeros(3 * n)
    lib().ora_collide(nx, ny, nz, ctypes.byref(p), _p(f),
                      _p(None if F is None else np.ascontiguousarray(F)),
                      _p(rho), _p(u))
    return f, rho, u


def propagate(shape, f):
    nx, ny, nz = shape
    out = np.empty_like(f)
    lib().ora_propagate(nx, ny, nz, _p(np.ascontiguousarray(f)), _p(out))
    return out


def equilibrium(shape, rho=None, u=None, rho_uniform=1.0):
    nx, ny, nz = shape
    f = np.zeros(19 * nx * ny * nz)
    lib().ora_equilibrium(nx, ny, nz, _p(rho), _p(u),
                          ctypes.c_double(rho_uniform), _p(f))
    return f


def grad27(shape, phi):
    nx, ny, nz = shape
    n = nx * ny * nz
    g = np.zeros(3 * n)
    d = np.zeros(n)
    lib().ora_grad27(nx, ny, nz, _p(np.ascontiguousarray(phi)), _p(g), _p(d))
    return g, d


def grad7(shape, phi):
    nx, ny, nz = shape
    n = nx * ny * nz
    g = np.zeros(3 * n)
    d = np.zeros(n)
    lib().ora_grad7(nx, ny, nz, _p(np.ascontiguousarray(phi)), _p(g), _p(d))
    return g, d


def mu(fe, phi, delsq):
    out = np.zeros(phi.size)
    lib().ora_mu(phi.size, ctypes.byref(fe), _p(phi), _p(delsq), _p(out))
    return out


def stress(fe, phi, grad, delsq):
    out = np.zeros(9 * phi.size)
    lib().ora_stress(phi.size, ctypes.byref(fe), _p(phi), _p(grad), _p(delsq),
                     _p(out))
    return out


def force(shape, fe, phi, grad, delsq):
    nx, ny, nz = shape
    F = np.zeros(3 * phi.size)
    lib().ora_force(nx, ny, nz, ctypes.byref(fe), _p(phi), _p(grad),
                    _p(delsq), _p(F))
    return F


def cahn_hilliard(shape, fe, phi, delsq, u):
    nx, ny, nz = shape
    out = np.ascontiguousarray(phi, dtype=np.float64).copy()
    lib().ora_cahn_hilliard(nx, ny, nz, ctypes.byref(fe), _p(out), _p(delsq),
                            _p(np.ascontiguousarray(u)))
    return out


def cahn_hilliard_ext(shape, fe, phi, delsq, u, csum=None, phi_init_sum=0.0):
    """The update with fe.conserve; returns (phi, csum)."""
    nx, ny, nz = shape
    out = np.ascontiguousarray(phi, dtype=np.float64).copy()
    cs = np.zeros(out.size) if csum is None else np.ascontiguousarray(csum, dtype=np.float64).copy()
    lib().ora_cahn_hilliard_ext(nx, ny, nz, ctypes.byref(fe), _p(out), _p(delsq),
                                _p(np.ascontiguousarray(u)), _p(cs), ctypes.c_double(phi_init_sum))
    return out, cs


def phi_init_sum(phi):
    a = np.ascontiguousarray(phi, dtype=np.float64)
    return lib().ora_phi_init_sum(a.size, _p(a))


def binary_steps_ext(shape, nsteps, lb, fe, f, phi, rho=None, u=None, csum=None,
                     phi_init_sum=0.0):
    """binary_steps with the conservation state; returns f, phi, rho, u, csum."""
    nx, ny, nz = shape
    n = nx * ny * nz
    f = np.ascontiguousarray(f, dtype=np.float64).copy()
    phi = np.ascontiguousarray(phi, dtype=np.float64).copy()
    rho = np.ones(n) if rho is None else rho.copy()
    u = np.zeros(3 * n) if u is None else u.copy()
    cs = np.zeros(n) if csum is None else np.ascontiguousarray(csum, dtype=np.float64).copy()
    lib().ora_binary_steps_ext(nx, ny, nz, nsteps, ctypes.byref(lb), ctypes.byref(fe), _p(f),
                               _p(phi), _p(rho), _p(u), _p(cs), ctypes.c_double(phi_init_sum))
    return f, phi, rho, u, cs


def binary_steps(shape, nsteps, lb, fe, f, phi, rho=None, u=None):
    nx, ny, nz = shape
    n = nx * ny * nz
    f = np.ascontiguousarray(f, dtype=np.float64).copy()
    phi = np.ascontiguousarray(phi, dtype=np.float64).copy()
    rho = np.ones(n) if rho is None else rho.copy()
    u = np.zeros(3 * n) if u is None else u.copy()
    lib().ora_binary_steps(nx, ny, nz, nsteps, ctypes.byref(lb),
                           ctypes.byref(fe), _p(f), _p(phi), _p(rho), _p(u))
    return f, phi, rho, u


def single_steps(shape, nsteps, lb, f):
    nx, ny, nz = shape
    n = nx * ny * nz
    f = np.ascontiguousarray(f, dtype=np.float64).copy()
    rho = np.zeros(n)
    u = np.zeros(3 * n)
    lib().ora_single_steps(nx, ny, nz, nsteps, ctypes.byref(lb), _p(f),
                           _p(rho), _p(u))
    return f, rho, u


def ns_uniform(state):
    s = (ctypes.c_uint32 * 4)(*state)
    r = lib().ora_ns_uniform(s)
    return r, list(s)


def spinodal(shape, seed, phi0, amp):
    nx, ny, nz = shape
    phi = np.zeros(nx * ny * nz)
    lib().ora_spinodal(nx, ny, nz, seed, ctypes.c_double(phi0),
                       ctypes.c_double(amp), _p(phi))
    return phi


def random_velocity(shape, seed, amp=1.0e-3):
    nx, ny, nz = shape
    u = np.zeros(3 * nx * ny * nz)
    lib().ora_random_velocity(nx, ny, nz, seed, ctypes.c_double(amp), _p(u))
    return u


def kahan_sum(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().ora_kahan_sum(a.size, _p(a))


def klein_sum(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().ora_klein_sum(a.size, _p(a))


# ---------------------------------------------------------------------------
# FP64 convolution checker (conv_oracle.c); channels-last arrays, periodic


def _conv_args(batch, nx, ny, nz, cin, cout):
    return [ctypes.c_int(v) for v in (batch, nx, ny, nz, cin, cout)]


def conv3_fwd(x, w, bias, shape, cin, cout, batch=1):
    nx, ny, nz = shape
    x = np.ascontiguousarray(x, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    b = None if bias is None else np.ascontiguousarray(bias, dtype=np.float64)
    y = np.zeros(batch * nx * ny * nz * cout)
    lib().ora_conv3_fwd(_p(x), _p(w), _p(b), _p(y), *_conv_args(batch, nx, ny, nz, cin, cout))
    return y


def conv3_bwd_data(dy, w, shape, cin, cout, batch=1):
    nx, ny, nz = shape
    dy = np.ascontiguousarray(dy, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    dx = np.zeros(batch * nx * ny * nz * cin)
    lib().ora_conv3_bwd_data(_p(dy), _p(w), _p(dx), *_conv_args(batch, nx, ny, nz, cin, cout))
    return dx


def conv3_bwd_weight(x, dy, shape, cin, cout, batch=1):
    nx, ny, nz = shape
    x = np.ascontiguousarray(x, dtype=np.float64)
    dy = np.ascontiguousarray(dy, dtype=np.float64)
    dw = np.zeros(27 * cout * cin)
    lib().ora_conv3_bwd_weight(_p(x), _p(dy), _p(dw), *_conv_args(batch, nx, ny, nz, cin, cout))
    return dw



# Lees-Edwards planes (le_oracle.c)

def _le_ck(rc):
    if rc != 0:
        raise ValueError("plane geometry refused (planes must divide nx and sit more "
                         "than nhalo sites from the x boundaries)")


def le_force(shape, fe, le, t, phi):
    """Gradient and flux-form force at physics time t; returns F (3 planes)."""
    nx, ny, nz = shape
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    F = np.zeros(3 * phi.size)
    _le_ck(lib().ora_le_force(nx, ny, nz, ctypes.byref(fe), ctypes.byref(le), t, _p(phi), _p(F)))
    return F


def le_cahn_hilliard(shape, fe, le, t, phi, u=None, csum=None, phi_init_sum=0.0):
    """Gradient and one Cahn-Hilliard update at time t; returns (phi, csum)."""
    nx, ny, nz = shape
    out = np.ascontiguousarray(phi, dtype=np.float64).copy()
    cs = np.zeros(out.size) if csum is None else np.ascontiguousarray(csum, dtype=np.float64).copy()
    uu = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
    _le_ck(lib().ora_le_cahn_hilliard(nx, ny, nz, ctypes.byref(fe), ctypes.byref(le), t,
                                      _p(out), _p(uu), _p(cs), phi_init_sum))
    return out, cs


def le_lb_bc(shape, le, timestep, f):
    nx, ny, nz = shape
    f = np.ascontiguousarray(f, dtype=np.float64).copy()
    _le_ck(lib().ora_le_lb_bc(nx, ny, nz, ctypes.byref(le), float(timestep), _p(f)))
    return f


def le_shear_profile(shape, le, rho0, eta):
    nx, ny, nz = shape
    f = np.zeros(19 * nx * ny * nz)
    _le_ck(lib().ora_le_shear_profile(nx, ny, nz, ctypes.byref(le), rho0, eta, _p(f)))
    return f


def le_binary_steps(shape, nsteps, lb, fe, le, f, phi, rho=None, u=None, csum=None,
                    phi_init_sum=0.0, t_start=0, t_current=0):
    """nsteps of the step loop with planes; returns f, phi, rho, u, csum."""
    nx, ny, nz = shape
    n = nx * ny * nz
    f = np.ascontiguousarray(f, dtype=np.float64).copy()
    phi = np.ascontiguousarray(phi, dtype=np.float64).copy()
    rho = np.ones(n) if rho is None else rho.copy()
    u = np.zeros(3 * n) if u is None else u.copy()
    cs = np.zeros(n) if csum is None else np.ascontiguousarray(csum, dtype=np.float64).copy()
    _le_ck(lib().ora_le_binary_steps(nx, ny, nz, nsteps, t_start, t_current, ctypes.byref(lb),
                                     ctypes.byref(fe), ctypes.byref(le), _p(f), _p(phi), _p(rho),
                                     _p(u), _p(cs), phi_init_sum))
    return f, phi, rho, u, cs
