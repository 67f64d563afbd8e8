"""Configuration files through the C ABI (lb_io_*, field_io_*, hydro_io_*):
a run that stops at step t, writes its configuration, and is restarted from
the files in a fresh object continues exactly as the uninterrupted run
(FP64, binary records), as the reference's restart does; ascii records hold
every value to 16 significant digits; missing or short files and I/O
decompositions other than one file fail loudly.

Byte-for-byte agreement of the files with the reference's Ludwig.exe is
tested through the host driver (tests/test_gpu_driver.py, iodrop chain).
"""

import os

import numpy as np
import pytest

import ludwig_amd as L

pytestmark = pytest.mark.gpu

SHAPE = (24, 20, 16)


def fluid(precision=64, order=1, stencil=7, io_format=L.IO_RECORD_BINARY, fatal=False):
    return L.BinaryFluid(SHAPE, a=-0.0625, b=0.0625, kappa=0.04, mobility=0.25, tau=0.8,
                         order=order, stencil=stencil, precision=precision, fatal=fatal,
                         io_format=io_format)


def state(b):
    rho, u, _ = b.hydro_fields()
    return b.f(), b.phi_field(), rho, u


@pytest.mark.parametrize("order,stencil", [(1, 7), (3, 7), (2, 27)])
def test_restart_continues_exactly(order, stencil, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    print(f"restart order {order} stencil {stencil}", flush=True)
    a = fluid(order=order, stencil=stencil)
    a.init_spinodal(8361235, phi0=0.05, amp=0.1)
    a.step(7)
    a.write_config(7)
    a.step(6)
    want = state(a)
    a.close()
    for fn in ("dist", "phi", "rho", "vel"):
        assert os.path.exists(f"{fn}-000000007.001-001") and os.path.exists(f"{fn}-metadata.001-001")
    b = fluid(order=order, stencil=stencil)
    b.init_spinodal(1, phi0=0.0, amp=0.2)      # overwritten by the files
    b.read_config(7)
    b.step(6)
    got = state(b)
    b.close()
    for name, g, w in zip(("f", "phi", "rho", "u"), got, want):
        bad = np.flatnonzero(g != w)
        assert bad.size == 0, f"{name}: {bad.size} sites differ after the restart"


def test_restart_fp32(tmp_path, monkeypatch):
    """FP32-shifted storage: the files hold doubles; phi, rho and u return
    exactly, f as (f - w rho0) rounded back to float, so the restarted run
    agrees to FP32 round-off."""
    monkeypatch.chdir(tmp_path)
    a = fluid(precision=32)
    a.init_spinodal(8361235, phi0=0.05, amp=0.1)
    a.step(5)
    phi5 = a.phi_field()
    a.write_config(5)
    a.step(5)
    want = state(a)
    a.close()
    b = fluid(precision=32)
    b.read_config(5)
    assert np.array_equal(b.phi_field(), phi5)
    b.step(5)
    got = state(b)
    b.close()
    for name, g, w, tol in zip(("f", "phi", "rho", "u"), got, want, (1e-7, 1e-6, 1e-7, 1e-9)):
        err = float(np.max(np.abs(g - w)))
        print(f"fp32 restart {name}: max abs diff {err:.3e}", flush=True)
        assert err <= tol, name


def test_ascii_records_round_trip(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    a = fluid(io_format=L.IO_RECORD_ASCII)
    a.init_spinodal(8361235, phi0=0.05, amp=0.1)
    a.step(4)
    want = state(a)
    a.write_config(4)
    a.close()
    line = open("vel-000000004.001-001").readline()
    assert len(line) == 3 * 23 + 1 and line.endswith("\n")
    with open("dist-000000004.001-001") as f:
        first = [f.readline() for _ in range(19)]
    assert all(len(x) == 24 for x in first)
    b = fluid(io_format=L.IO_RECORD_ASCII)
    b.read_config(4)
    got = state(b)
    b.close()
    for name, g, w in zip(("f", "phi", "rho", "u"), got, want):
        # " %22.15e" keeps 16 significant digits
        np.testing.assert_allclose(g, w, rtol=6e-16, atol=1e-300, err_msg=name)


def test_missing_and_short_files_fail(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    a = fluid()
    a.init_spinodal(8361235)
    with pytest.raises(RuntimeError, match="dist-000000003.001-001"):
        a.read_config(3)
    a.write_config(3)
    with open("phi-000000003.001-001", "r+b") as f:
        f.truncate(100)
    with pytest.raises(RuntimeError, match="holds 100 bytes"):
        a.read_config(3)
    a.close()


def test_io_grid_other_than_one_file_is_refused():
    L.lib().ldw_error_mode_set(0)
    opts = L.lib().lb_data_options_default()
    assert opts.iodata.output.iorformat == L.IO_RECORD_BINARY
    assert list(opts.iodata.output.iogrid) == [1, 1, 1]
    b = fluid()
    import ctypes
    opts.iodata.output.iogrid[0] = 2
    lb = ctypes.c_void_p()
    rc = L.lib().lb_data_create(b.pe, b.cs, ctypes.byref(opts), ctypes.byref(lb))
    assert rc != 0 and "i/o grid 2 1 1" in L.lib().ldw_last_error().decode()
    b.close()
