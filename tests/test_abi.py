"""CPU checks of the C ABI: libludwig_hip.so loads and exports every function
that include/ludwig_hip.h declares (no device calls are made here)."""

import ctypes
import subprocess

import pytest

import ludwig_amd


def test_header_declares_the_reference_entry_points():
    names = set(ludwig_amd.header_functions())
    for n in ["lb_data_create", "lb_collide", "lb_propagation", "lb_halo", "lb_memcpy",
              "field_create", "field_halo", "field_grad_compute", "hydro_create",
              "hydro_f_zero", "hydro_u_zero", "fe_symm_create", "fe_symm_param_set",
              "pth_create", "phi_force_calculation", "phi_ch_create", "phi_cahn_hilliard",
              "field_phi_init_spinodal", "lb_init_rest_f", "lb_1st_moment_equilib_set"]:
        assert n in names, n


def test_library_exports_every_declared_function():
    import os
    if not os.path.exists(ludwig_amd.LIB_PATH):
        pytest.fail("libludwig_hip.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(ludwig_amd.LIB_PATH)
    missing = [n for n in ludwig_amd.header_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", ludwig_amd.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for n in ludwig_amd.header_functions():
        assert n in exported, n


def test_library_has_gfx950_code_object():
    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", ludwig_amd.LIB_PATH],
                         capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("roc-obj-ls unavailable")
    assert "gfx950" in out.stdout
