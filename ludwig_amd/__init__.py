"""ludwig_amd -- Python host side of libludwig_hip.so (the MI355X D3Q19
binary-fluid step behind Ludwig's C object API).

The compute lives in ``libludwig_hip.so`` (HIP kernels for gfx950 plus the
C ABI declared in ``include/ludwig_hip.h``).  This module binds that ABI with
ctypes and offers two conveniences that build the reference's object graph
the way ``ludwig_rt`` does (ludwig.c:202-429):

* :class:`BinaryFluid` -- symmetric free energy + Cahn-Hilliard + D3Q19 MRT
  (the ``free_energy symmetric`` branch, ludwig.c:1190-1262);
* :class:`SingleFluid` -- ``free_energy none`` (ludwig.c:1175-1189).

There is no CPU fallback: loading fails loudly when the library is not
built, and object creation fails when no GPU is present.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libludwig_hip.so")
HEADER = os.path.join(REPO, "include", "ludwig_hip.h")

# enums of the reference (lb_data_options.h:20, fe_force_method.h:20)
M10, BGK, TRT = 0, 1, 2
H2D, D2H = 1, 2
FORCE_NO_FORCE, FORCE_STRESS_DIVERGENCE, FORCE_PHI_GRADMU = 1, 2, 3
LE_SHEAR_TYPE_STEADY = 1                    # lees_edwards_options.h:26

_lib = None


IO_RECORD_ASCII, IO_RECORD_BINARY = 1, 2


class IoOptions(ctypes.Structure):
    """io_options_t (io_options.h:51)."""
    _fields_ = [("mode", ctypes.c_int), ("iorformat", ctypes.c_int),
                ("metadata_version", ctypes.c_int), ("report", ctypes.c_int),
                ("asynchronous", ctypes.c_int), ("compression_levl", ctypes.c_int),
                ("iogrid", ctypes.c_int * 3)]


class IoInfoArgs(ctypes.Structure):
    """io_info_args_t (io_info_args.h:30)."""
    _fields_ = [("input", IoOptions), ("output", IoOptions), ("grid", ctypes.c_int * 3),
                ("iofreq", ctypes.c_int)]


class LbDataOptions(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int), ("nvel", ctypes.c_int), ("ndist", ctypes.c_int),
                ("nrelax", ctypes.c_int), ("halo", ctypes.c_int),
                ("reportimbalance", ctypes.c_int), ("usefirsttouch", ctypes.c_int),
                ("iodata", IoInfoArgs)]


class FieldOptions(ctypes.Structure):
    _fields_ = [("ndata", ctypes.c_int), ("nhcomm", ctypes.c_int), ("haloscheme", ctypes.c_int),
                ("haloverbose", ctypes.c_int), ("usefirsttouch", ctypes.c_int),
                ("iodata", IoInfoArgs)]


class HydroOptions(ctypes.Structure):
    _fields_ = [("nhcomm", ctypes.c_int), ("rho", FieldOptions), ("u", FieldOptions),
                ("force", FieldOptions), ("eta", FieldOptions)]


class IoEvent(ctypes.Structure):
    """io_event_t (io_event.h:35)."""
    _fields_ = [("extra_name", ctypes.c_char_p), ("extra_json", ctypes.c_void_p),
                ("time", ctypes.c_double * 6)]


class MapOptions(ctypes.Structure):
    _fields_ = [("ndata", ctypes.c_int), ("is_porous_media", ctypes.c_int)]


class LeOptions(ctypes.Structure):
    _fields_ = [("nplanes", ctypes.c_int), ("type", ctypes.c_int), ("period", ctypes.c_int),
                ("nt0", ctypes.c_int), ("uy", ctypes.c_double)]


class FeSymmParam(ctypes.Structure):
    _fields_ = [("a", ctypes.c_double), ("b", ctypes.c_double), ("kappa", ctypes.c_double),
                ("c", ctypes.c_double), ("h", ctypes.c_double)]


class PhiChInfo(ctypes.Structure):
    _fields_ = [("conserve", ctypes.c_int), ("noise", ctypes.c_int)]


def header_functions():
    """Names of the functions declared in include/ludwig_hip.h."""
    import re
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = "\n".join(l for l in text.splitlines() if "typedef" not in l)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                       text, flags=re.M)
    return sorted(set(n for n in names if n not in ("typedef",)))


def lib():
    """Load libludwig_hip.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `make -C ludwig_amd` "
                           "(or __graft_entry__.build()); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    pvp = ctypes.POINTER(ctypes.c_void_p)
    dp = ctypes.c_void_p
    sig = {
        "ldw_error_mode_set": [ctypes.c_int], "ldw_precision_set": [ctypes.c_int],
        "ldw_device_synchronize": [],
        "pe_create": [ctypes.c_int, ctypes.c_int, pvp], "pe_free": [vp],
        "cs_create": [vp, pvp], "cs_free": [vp], "cs_init": [vp],
        "cs_ntotal_set": [ctypes.POINTER(ctypes.c_int)], "cs_nhalo_set": [vp, ctypes.c_int],
        "cs_index": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int],
        "physics_create": [vp, pvp], "physics_free": [vp],
        "physics_rho0_set": [vp, ctypes.c_double], "physics_eta_shear_set": [vp, ctypes.c_double],
        "physics_eta_bulk_set": [vp, ctypes.c_double], "physics_mobility_set": [vp, ctypes.c_double],
        "physics_fbody_set": [vp, ctypes.POINTER(ctypes.c_double)],
        "physics_grad_mu_set": [vp, ctypes.POINTER(ctypes.c_double)],
        "lb_data_create": [vp, vp, ctypes.POINTER(LbDataOptions), pvp], "lb_free": [vp],
        "lb_memcpy": [vp, ctypes.c_int], "lb_halo": [vp], "lb_init_rest_f": [vp, ctypes.c_double],
        "lb_collision_relaxation_set": [vp, ctypes.c_int],
        "lb_collide": [vp, vp, vp, vp, vp, vp], "lb_propagation": [vp],
        "lees_edw_create": [vp, vp, ctypes.POINTER(LeOptions), pvp], "lees_edw_free": [vp],
        "lees_edw_nplane_total": [vp], "lees_edw_plane_location": [vp, ctypes.c_int],
        "lees_edw_shear_rate": [vp, ctypes.POINTER(ctypes.c_double)],
        "lb_data_apply_le_boundary_conditions": [vp, vp], "lb_le_init_shear_profile": [vp, vp],
        "physics_control_init_time": [vp, ctypes.c_int, ctypes.c_int],
        "physics_control_next_step": [vp], "physics_control_timestep": [vp],
        "physics_control_time": [vp, ctypes.POINTER(ctypes.c_double)],
        "map_create": [vp, vp, ctypes.POINTER(MapOptions), pvp], "map_free": [pvp],
        "wall_create": [vp, vp, vp, vp, pvp], "wall_free": [vp],
        "field_create": [vp, vp, vp, ctypes.c_char_p, ctypes.POINTER(FieldOptions), pvp],
        "field_free": [vp], "field_memcpy": [vp, ctypes.c_int], "field_halo": [vp],
        "field_grad_create": [vp, vp, ctypes.c_int, pvp], "field_grad_free": [vp],
        "field_grad_compute": [vp], "field_grad_memcpy": [vp, ctypes.c_int],
        "field_grad_set": [vp, vp, vp],
        "field_grad_scalar_grad": [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
        "field_grad_scalar_delsq": [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
        "fe_symm_mu": [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
        "fe_symm_fed": [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
        "fe_symm_str": [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
        "hydro_create": [vp, vp, vp, ctypes.POINTER(HydroOptions), pvp], "hydro_free": [vp],
        "hydro_f_zero": [vp, ctypes.POINTER(ctypes.c_double)],
        "hydro_u_zero": [vp, ctypes.POINTER(ctypes.c_double)],
        "hydro_memcpy": [vp, ctypes.c_int],
        "fe_symm_create": [vp, vp, vp, vp, pvp], "fe_symm_free": [vp],
        "fe_symm_param_set": [vp, FeSymmParam],
        "pth_create": [vp, vp, ctypes.c_int, pvp], "pth_free": [vp],
        "phi_force_calculation": [vp, vp, vp, vp, vp, vp, vp, vp, vp],
        "phi_ch_create": [vp, vp, vp, ctypes.POINTER(PhiChInfo), pvp], "phi_ch_free": [vp],
        "phi_cahn_hilliard": [vp, vp, vp, vp, vp, vp],
        "advection_order_set": [ctypes.c_int],
        "ldw_conserve2_site_order": [ctypes.c_int],
        "cahn_hilliard_stats_time0": [vp, vp, vp],
        "field_phi_init_spinodal": [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double],
        "ldw_lb_init_random_velocity": [vp, ctypes.c_int, ctypes.c_double],
        "ldw_step_binary": [vp, vp, vp, vp, vp, vp, vp, ctypes.c_int],
        "ldw_step_single": [vp, vp, ctypes.c_int],
        "ldw_conserved_sums": [vp, vp, ctypes.POINTER(ctypes.c_double)],
        "ldw_timer_enable": [ctypes.c_int], "ldw_timer_reset": [],
        "ldw_timer_read": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)],
        "ldw_stream_mark": [ctypes.c_int],
        "ldw_stream_elapsed": [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
        "ldw_lb_f_read": [vp, dp], "ldw_lb_f_write": [vp, dp],
        "ldw_field_read": [vp, dp], "ldw_field_write": [vp, dp],
        "ldw_hydro_read": [vp, dp, dp, dp], "ldw_hydro_write": [vp, dp, dp, dp],
        "ldw_copy_calibration": [ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
        "ldw_device_alloc": [ctypes.c_longlong, pvp], "ldw_device_free": [vp],
        "ldw_memcpy": [vp, vp, ctypes.c_longlong, ctypes.c_int],
        # surrogate network kernels (device pointers)
        "ldw_conv3d_fwd": [vp, vp, vp, vp] + [ctypes.c_int] * 7,
        "ldw_conv3d_fwd_silu": [vp, vp, vp, vp, vp] + [ctypes.c_int] * 6,
        "ldw_conv3d_bwd_data": [vp, vp, vp] + [ctypes.c_int] * 7,
        "ldw_conv3d_bwd_weight": [vp, vp, vp] + [ctypes.c_int] * 7,
        "ldw_conv3d_weights_prep": [vp, vp, vp, ctypes.c_int, ctypes.c_int],
        "ldw_nn_silu_fwd": [vp, vp, ctypes.c_longlong],
        "ldw_nn_silu_bwd": [vp, vp, vp, vp, ctypes.c_longlong],
        "ldw_nn_bias_grad": [vp, vp, ctypes.c_longlong, ctypes.c_int, ctypes.c_int],
        "ldw_nn_pool2_fwd": [vp, vp] + [ctypes.c_int] * 5,
        "ldw_nn_pool2_bwd": [vp, vp] + [ctypes.c_int] * 6,
        "ldw_nn_up2_fwd": [vp, vp] + [ctypes.c_int] * 5,
        "ldw_nn_up2_bwd": [vp, vp] + [ctypes.c_int] * 5,
        "ldw_nn_concat": [vp, ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_longlong],
        "ldw_nn_up2_concat": [vp, ctypes.c_int, vp, ctypes.c_int, vp] + [ctypes.c_int] * 4,
        "ldw_nn_split": [vp, ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_longlong, ctypes.c_int,
                         ctypes.c_int],
        "ldw_nn_to_bf16": [vp, vp, ctypes.c_longlong],
        "ldw_nn_encode": [vp, ctypes.c_longlong, ctypes.POINTER(ctypes.c_float), vp, ctypes.c_int],
        "ldw_nn_target": [vp, vp, ctypes.c_longlong, vp, ctypes.c_float, ctypes.c_float],
        "ldw_nn_decode": [vp, vp, ctypes.c_int, ctypes.c_longlong, vp] + [ctypes.c_float] * 4,
        "ldw_nn_mse": [vp, ctypes.c_int, vp, ctypes.c_longlong, vp, vp, ctypes.POINTER(ctypes.c_float),
                       ctypes.POINTER(ctypes.c_double)],
        "ldw_nn_adam": [vp, vp, vp, vp, ctypes.c_longlong] + [ctypes.c_float] * 6 + [ctypes.c_int],
        "ldw_nn_sumsq": [vp, ctypes.c_longlong, ctypes.POINTER(ctypes.c_double)],
        "ldw_nn_scale": [vp, ctypes.c_longlong, ctypes.c_float],
        "ldw_nn_sum": [vp, ctypes.c_longlong, ctypes.POINTER(ctypes.c_double)],
        "ldw_nn_shift": [vp, ctypes.c_longlong, ctypes.c_float],
        "ldw_nn_restore_sums": [vp, vp, ctypes.c_longlong, ctypes.c_int],
        "ldw_nn_restore_sums_noalloc": [vp, vp, ctypes.c_longlong, ctypes.c_int],
        "ldw_nn_cubic": [vp, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.cs_ntotal_set.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    for name in ("lb_io_write", "lb_io_read", "field_io_write", "field_io_read",
                 "hydro_io_write", "hydro_io_read"):
        f = getattr(L, name)
        f.argtypes = [vp, ctypes.c_int, ctypes.POINTER(IoEvent)]
        f.restype = ctypes.c_int
    L.field_phi_init_drop.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int]
    L.field_phi_init_drop.restype = ctypes.c_int
    L.lb_data_options_default.argtypes = []
    L.lb_data_options_default.restype = LbDataOptions
    L.field_options_ndata_nhalo.argtypes = [ctypes.c_int, ctypes.c_int]
    L.field_options_ndata_nhalo.restype = FieldOptions
    L.hydro_options_default.argtypes = []
    L.hydro_options_default.restype = HydroOptions
    L.ldw_last_error.restype = ctypes.c_char_p
    L.ldw_last_error.argtypes = []
    L.ldw_precision.restype = ctypes.c_int
    _lib = L
    return L


def _ck(rc, what):
    if rc != 0:
        msg = lib().ldw_last_error().decode()
        raise RuntimeError(f"{what} failed: {msg}")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class DeviceArray:
    """A 1-D array in device memory allocated through libludwig_hip (the
    library's HIP runtime), with host copies through numpy."""

    def __init__(self, n, dtype=np.float32):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        p = ctypes.c_void_p()
        _ck(lib().ldw_device_alloc(self.nbytes, ctypes.byref(p)), "ldw_device_alloc")
        self.ptr = p.value or 0

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype if dtype is not None else np.asarray(a).dtype)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n, (a.size, self.n)
        _ck(lib().ldw_memcpy(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p),
                             self.nbytes, H2D), "ldw_memcpy")

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        _ck(lib().ldw_memcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                             self.nbytes, D2H), "ldw_memcpy")
        return out

    def free(self):
        if self.ptr and _lib is not None:
            _lib.ldw_device_free(ctypes.c_void_p(self.ptr))
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def set_precision(bits):
    _ck(lib().ldw_precision_set(bits), "ldw_precision_set")


class _Base:
    """The reference's pe/cs/physics/lb/hydro/map objects for an
    nx*ny*nz periodic lattice."""

    def __init__(self, shape, nhalo, tau=None, eta=None, eta_bulk=None, rho0=1.0,
                 relax=M10, fbody=(0.0, 0.0, 0.0), precision=64, fatal=False,
                 io_format=IO_RECORD_BINARY, le_planes=0, le_uy=0.0, le_nt0=0):
        """io_format: record format of configuration files, IO_RECORD_BINARY
        or IO_RECORD_ASCII, or an (input, output) pair.  le_planes, le_uy,
        le_nt0: Lees-Edwards planes (N_LE_plane, LE_plane_vel,
        LE_time_offset; steady shear, FP64)."""
        L = lib()
        L.ldw_error_mode_set(1 if fatal else 0)
        set_precision(precision)
        self.precision = precision
        self.shape = tuple(int(s) for s in shape)
        self.nsite = int(np.prod(self.shape))
        if eta is None:
            eta = (tau - 0.5) / 3.0
        self.eta = eta
        self.eta_bulk = eta if eta_bulk is None else eta_bulk
        vp = ctypes.c_void_p
        self.pe, self.cs, self.phys = vp(), vp(), vp()
        _ck(L.pe_create(0, 0, ctypes.byref(self.pe)), "pe_create")
        _ck(L.physics_create(self.pe, ctypes.byref(self.phys)), "physics_create")
        _ck(L.physics_rho0_set(self.phys, rho0), "physics_rho0_set")
        _ck(L.physics_eta_shear_set(self.phys, eta), "physics_eta_shear_set")
        _ck(L.physics_eta_bulk_set(self.phys, self.eta_bulk), "physics_eta_bulk_set")
        fb = (ctypes.c_double * 3)(*fbody)
        _ck(L.physics_fbody_set(self.phys, fb), "physics_fbody_set")
        _ck(L.cs_create(self.pe, ctypes.byref(self.cs)), "cs_create")
        _ck(L.cs_nhalo_set(self.cs, nhalo), "cs_nhalo_set")
        nt = (ctypes.c_int * 3)(*self.shape)
        _ck(L.cs_ntotal_set(self.cs, nt), "cs_ntotal_set")
        _ck(L.cs_init(self.cs), "cs_init")
        self.le = vp()
        leo = LeOptions(le_planes, LE_SHEAR_TYPE_STEADY if le_planes else 0, 0, le_nt0, le_uy)
        _ck(L.lees_edw_create(self.pe, self.cs, ctypes.byref(leo), ctypes.byref(self.le)),
            "lees_edw_create")
        self.le_planes = le_planes
        opts = L.lb_data_options_default()
        opts.nrelax = relax
        io = io_format if isinstance(io_format, tuple) else (io_format, io_format)
        self.io_format = io
        opts.iodata.input.iorformat, opts.iodata.output.iorformat = io
        self.lb = vp()
        _ck(L.lb_data_create(self.pe, self.cs, ctypes.byref(opts), ctypes.byref(self.lb)),
            "lb_data_create")
        _ck(L.lb_collision_relaxation_set(self.lb, relax), "lb_collision_relaxation_set")
        self.hydro = vp()
        ho = L.hydro_options_default()
        for fo in (ho.rho, ho.u):
            fo.iodata.input.iorformat, fo.iodata.output.iorformat = io
        _ck(L.hydro_create(self.pe, self.cs, self.le, ctypes.byref(ho),
                           ctypes.byref(self.hydro)), "hydro_create")
        self.map = vp()
        _ck(L.map_create(self.pe, self.cs, ctypes.byref(MapOptions(0, 0)), ctypes.byref(self.map)),
            "map_create")

    # state access (interior, device layout, float64)
    def f(self):
        out = np.empty(19 * self.nsite)
        _ck(lib().ldw_lb_f_read(self.lb, _ptr(out)), "ldw_lb_f_read")
        return out

    def set_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.size == 19 * self.nsite
        _ck(lib().ldw_lb_f_write(self.lb, _ptr(f)), "ldw_lb_f_write")

    def hydro_fields(self):
        rho = np.empty(self.nsite)
        u = np.empty(3 * self.nsite)
        force = np.empty(3 * self.nsite)
        _ck(lib().ldw_hydro_read(self.hydro, _ptr(rho), _ptr(u), _ptr(force)), "ldw_hydro_read")
        return rho, u, force

    def set_hydro(self, rho=None, u=None, force=None):
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        rho, u, force = c(rho), c(u), c(force)
        _ck(lib().ldw_hydro_write(self.hydro, _ptr(rho), _ptr(u), _ptr(force)), "ldw_hydro_write")

    def conserved_sums(self, phi=None):
        out = (ctypes.c_double * 2)()
        _ck(lib().ldw_conserved_sums(self.lb, phi, out), "ldw_conserved_sums")
        return out[0], out[1]

    def synchronize(self):
        _ck(lib().ldw_device_synchronize(), "ldw_device_synchronize")

    # the reference's per-call API
    def collide(self, fe=None):
        _ck(lib().lb_collide(self.lb, self.hydro, self.map, None, fe, None), "lb_collide")

    def propagate(self):
        _ck(lib().lb_halo(self.lb), "lb_halo")
        _ck(lib().lb_propagation(self.lb), "lb_propagation")

    def init_rest(self, rho0=1.0):
        _ck(lib().lb_init_rest_f(self.lb, rho0), "lb_init_rest_f")

    def le_boundary_conditions(self):
        """lb_data_apply_le_boundary_conditions, between collide and propagate."""
        _ck(lib().lb_data_apply_le_boundary_conditions(self.lb, self.le),
            "lb_data_apply_le_boundary_conditions")

    def le_init_shear_profile(self):
        _ck(lib().lb_le_init_shear_profile(self.lb, self.le), "lb_le_init_shear_profile")

    # the time step counter the Lees-Edwards planes move with
    def next_step(self, n=1):
        for _ in range(n):
            lib().physics_control_next_step(self.phys)

    def time(self):
        t = ctypes.c_double()
        _ck(lib().physics_control_time(self.phys, ctypes.byref(t)), "physics_control_time")
        return t.value

    # configuration files of time step t in the working directory
    # (lb_io_write / hydro_io_write; lb_io_read / hydro_io_read)
    def write_config(self, t):
        ev = IoEvent()
        _ck(lib().lb_io_write(self.lb, t, ctypes.byref(ev)), "lb_io_write")
        _ck(lib().hydro_io_write(self.hydro, t, ctypes.byref(ev)), "hydro_io_write")

    def read_config(self, t):
        ev = IoEvent()
        _ck(lib().lb_io_read(self.lb, t, ctypes.byref(ev)), "lb_io_read")
        _ck(lib().hydro_io_read(self.hydro, t, ctypes.byref(ev)), "hydro_io_read")

    def f_zero(self, fzero=(0.0, 0.0, 0.0)):
        _ck(lib().hydro_f_zero(self.hydro, (ctypes.c_double * 3)(*fzero)), "hydro_f_zero")

    def u_zero(self, uzero=(0.0, 0.0, 0.0)):
        _ck(lib().hydro_u_zero(self.hydro, (ctypes.c_double * 3)(*uzero)), "hydro_u_zero")

    def close(self):
        L = _lib
        if L is None or self.lb is None:
            return
        L.ldw_device_synchronize()
        L.map_free(ctypes.byref(self.map))
        L.hydro_free(self.hydro)
        L.lb_free(self.lb)
        L.lees_edw_free(self.le)
        L.cs_free(self.cs)
        L.physics_free(self.phys)
        L.pe_free(self.pe)
        self.lb = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SingleFluid(_Base):
    """``free_energy none``: plain D3Q19 MRT fluid (nhalo 1)."""

    def __init__(self, shape, **kw):
        super().__init__(shape, 1, **kw)

    def init_random_velocity(self, seed, amp=1.0e-3):
        _ck(lib().ldw_lb_init_random_velocity(self.lb, seed, amp), "ldw_lb_init_random_velocity")

    def step(self, nsteps=1):
        _ck(lib().ldw_step_single(self.lb, self.hydro, nsteps), "ldw_step_single")


class BinaryFluid(_Base):
    """``free_energy symmetric`` with the finite-difference Cahn-Hilliard
    solver and the 7-point gradient (nhalo 2)."""

    def __init__(self, shape, a=-0.0625, b=0.0625, kappa=0.04, mobility=0.25,
                 order=1, force_method=FORCE_STRESS_DIVERGENCE, stencil=7, conserve=0,
                 grad_mu=(0.0, 0.0, 0.0), **kw):
        """conserve: cahn_hilliard_options_conserve 0, 1 (compensated
        update) or 2 (global subtraction; the reference sum is taken by
        the init_* and read_config methods).  order: advection order 1-5.
        grad_mu: the external chemical potential gradient (input key
        grad_mu)."""
        super().__init__(shape, 2, **kw)
        L = lib()
        vp = ctypes.c_void_p
        _ck(L.physics_mobility_set(self.phys, mobility), "physics_mobility_set")
        _ck(L.physics_grad_mu_set(self.phys, (ctypes.c_double * 3)(*grad_mu)), "physics_grad_mu_set")
        _ck(L.advection_order_set(order), "advection_order_set")
        self.order = order
        self.phi = vp()
        fo = L.field_options_ndata_nhalo(1, 2)
        fo.iodata.input.iorformat, fo.iodata.output.iorformat = self.io_format
        _ck(L.field_create(self.pe, self.cs, self.le, b"phi", ctypes.byref(fo), ctypes.byref(self.phi)),
            "field_create")
        self.dphi = vp()
        _ck(L.field_grad_create(self.pe, self.phi, 2, ctypes.byref(self.dphi)), "field_grad_create")
        if stencil not in (7, 27):
            raise ValueError(f"stencil {stencil}: 7 (3d_7pt_fluid) or 27 (3d_27pt_fluid)")
        d2 = L.grad_3d_27pt_fluid_d2 if stencil == 27 else L.grad_3d_7pt_fluid_d2
        _ck(L.field_grad_set(self.dphi, ctypes.cast(d2, ctypes.c_void_p), None), "field_grad_set")
        self.fe = vp()
        _ck(L.fe_symm_create(self.pe, self.cs, self.phi, self.dphi, ctypes.byref(self.fe)),
            "fe_symm_create")
        _ck(L.fe_symm_param_set(self.fe, FeSymmParam(a, b, kappa, 0.0, 0.0)), "fe_symm_param_set")
        self.pth = vp()
        _ck(L.pth_create(self.pe, self.cs, force_method, ctypes.byref(self.pth)), "pth_create")
        self.pch = vp()
        self.conserve = conserve
        _ck(L.phi_ch_create(self.pe, self.cs, self.le, ctypes.byref(PhiChInfo(conserve, 0)),
                            ctypes.byref(self.pch)), "phi_ch_create")
        self.wall = vp()
        _ck(L.wall_create(self.pe, self.cs, self.map, self.lb, ctypes.byref(self.wall)), "wall_create")

    def init_spinodal(self, seed, phi0=0.0, amp=0.1, rho0=1.0):
        _ck(lib().field_phi_init_spinodal(self.phi, seed, phi0, amp), "field_phi_init_spinodal")
        _ck(lib().lb_init_rest_f(self.lb, rho0), "lb_init_rest_f")
        self._reference_sum()

    def init_drop(self, radius, xi, phistar=1.0, centred=False, rho0=1.0):
        _ck(lib().field_phi_init_drop(self.phi, xi, radius, phistar, 1 if centred else 0),
            "field_phi_init_drop")
        _ck(lib().lb_init_rest_f(self.lb, rho0), "lb_init_rest_f")
        self._reference_sum()

    def _reference_sum(self):
        # ludwig.c:420-425: conserve 2 needs the initial sum
        if self.conserve == 2:
            _ck(lib().cahn_hilliard_stats_time0(self.pch, self.phi, self.map),
                "cahn_hilliard_stats_time0")

    def write_config(self, t):
        """dist, phi, rho and vel files of time step t, as a configuration
        step of the reference writes them (ludwig.c:868-962)."""
        super().write_config(t)
        _ck(lib().field_io_write(self.phi, t, ctypes.byref(IoEvent())), "field_io_write")

    def read_config(self, t):
        """The restart read of ludwig.c:313-364: dist, phi, rho and vel."""
        ev = IoEvent()
        _ck(lib().lb_io_read(self.lb, t, ctypes.byref(ev)), "lb_io_read")
        _ck(lib().field_io_read(self.phi, t, ctypes.byref(ev)), "field_io_read")
        _ck(lib().hydro_io_read(self.hydro, t, ctypes.byref(ev)), "hydro_io_read")
        self._reference_sum()

    def step(self, nsteps=1):
        _ck(lib().advection_order_set(self.order), "advection_order_set")
        _ck(lib().ldw_step_binary(self.lb, self.hydro, self.phi, self.dphi, self.fe, self.pth,
                                  self.pch, nsteps), "ldw_step_binary")

    # individual stages, as the reference's time loop calls them
    def gradient_stage(self):
        _ck(lib().field_halo(self.phi), "field_halo")
        _ck(lib().field_grad_compute(self.dphi), "field_grad_compute")

    def force_stage(self):
        _ck(lib().field_halo(self.phi), "field_halo")
        _ck(lib().field_grad_compute(self.dphi), "field_grad_compute")
        _ck(lib().phi_force_calculation(self.pe, self.cs, self.le, self.wall, self.pth, self.fe,
                                        self.map, self.phi, self.hydro), "phi_force_calculation")

    def cahn_hilliard_stage(self, with_hydro=True):
        _ck(lib().phi_cahn_hilliard(self.pch, self.fe, self.phi,
                                    self.hydro if with_hydro else None, self.map, None),
            "phi_cahn_hilliard")

    def gradient_readback(self):
        """The reference's host read-back after field_grad_compute():
        field_grad_memcpy(D2H), then field_grad_scalar_grad/_delsq and
        fe_symm_mu/_fed/_str at every interior site (cs_index addressing).
        Returns grad (3, nsite), delsq, mu, fed (nsite) and str (9, nsite),
        sites z fastest.  One call per site: for small lattices."""
        L = lib()
        _ck(L.field_memcpy(self.phi, D2H), "field_memcpy")
        _ck(L.field_grad_memcpy(self.dphi, D2H), "field_grad_memcpy")
        nx, ny, nz = self.shape
        grad = np.empty((3, self.nsite))
        out = {k: np.empty(self.nsite) for k in ("delsq", "mu", "fed")}
        st = np.empty((9, self.nsite))
        g3 = (ctypes.c_double * 3)()
        d = ctypes.c_double()
        s9 = (ctypes.c_double * 9)()
        s = 0
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    idx = L.cs_index(self.cs, i + 1, j + 1, k + 1)
                    _ck(L.field_grad_scalar_grad(self.dphi, idx, g3), "field_grad_scalar_grad")
                    grad[:, s] = g3[:]
                    for name, fn in (("delsq", L.field_grad_scalar_delsq), ("mu", L.fe_symm_mu),
                                     ("fed", L.fe_symm_fed)):
                        _ck(fn(self.dphi if name == "delsq" else self.fe, idx, ctypes.byref(d)), name)
                        out[name][s] = d.value
                    _ck(L.fe_symm_str(self.fe, idx, s9), "fe_symm_str")
                    st[:, s] = s9[:]
                    s += 1
        return grad, out["delsq"], out["mu"], out["fed"], st

    def phi_field(self):
        out = np.empty(self.nsite)
        _ck(lib().ldw_field_read(self.phi, _ptr(out)), "ldw_field_read")
        return out

    def set_phi(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        _ck(lib().ldw_field_write(self.phi, _ptr(phi)), "ldw_field_write")

    def conserved_sums(self):
        return super().conserved_sums(self.phi)

    def close(self):
        L = _lib
        if L is not None and getattr(self, "lb", None) is not None:
            L.ldw_device_synchronize()
            L.wall_free(self.wall)
            L.phi_ch_free(self.pch)
            L.pth_free(self.pth)
            L.fe_symm_free(self.fe)
            L.field_grad_free(self.dphi)
            L.field_free(self.phi)
        super().close()
