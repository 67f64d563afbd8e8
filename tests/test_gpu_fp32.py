"""FP32-shifted storage (f_i - w_i rho0 in FP32, FP32 arithmetic) against the
FP64 path of the same library and the reference's golden vectors.  There is
no reference counterpart; the bar is the error stated per field here."""

import numpy as np
import pytest

import ludwig_amd as L

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def test_fp32_collide_against_golden(golden):
    g = golden("collide_m10")
    n = int(g["n"])
    s = L.SingleFluid((n, n, n), tau=0.8, precision=32)
    s.set_f(g["f_in"])
    s.set_hydro(force=g["force"])
    s.u_zero()
    s.collide()
    f = s.f()
    rho, u, _ = s.hydro_fields()
    # f - w rho0 is O(1e-2) and stored in FP32: absolute error ~1e-9
    assert np.max(np.abs(f - g["f_out"])) < 5e-8
    assert np.max(np.abs(rho - g["rho"])) < 5e-7
    assert rel_rms(u, g["u"]) < 1e-5
    s.close()


@pytest.mark.parametrize("n,steps,phi_tol", [(32, 100, 1e-6), (64, 300, 1e-6)])
def test_fp32_binary_against_fp64(n, steps, phi_tol):
    out = {}
    for prec in (64, 32):
        b = L.BinaryFluid((n, n, n), tau=1.0, mobility=0.25, precision=prec)
        b.init_spinodal(2024, 0.0, 0.1)
        s0 = b.conserved_sums()
        b.step(steps)
        rho, u, _ = b.hydro_fields()
        out[prec] = (b.phi_field(), rho, u, s0, b.conserved_sums())
        b.close()
    phi64, rho64, u64 = out[64][:3]
    phi32, rho32, u32 = out[32][:3]
    print(f"n={n} steps={steps}: phi rel-rms {rel_rms(phi32, phi64):.2e}, "
          f"u rel-rms {rel_rms(u32, u64):.2e}")
    assert rel_rms(phi32, phi64) < phi_tol
    assert rel_rms(u32, u64) < 2e-5
    assert np.max(np.abs(rho32 - rho64)) < 1e-6
    # conservation from compensated FP64 device sums
    (r0, p0), (r1, p1) = out[32][3], out[32][4]
    assert abs(r1 - r0) / r0 < 1e-9
    assert abs(p1 - p0) / max(abs(p0), 1.0) < 1e-5


def test_fp32_single_fluid_c1_against_fp64():
    n = 32
    res = {}
    for prec in (64, 32):
        s = L.SingleFluid((n, n, n), tau=0.8, precision=prec)
        s.init_random_velocity(12345)
        s.step(100)
        rho, u, _ = s.hydro_fields()
        res[prec] = (s.f(), rho, u)
        s.close()
    assert rel_rms(res[32][2], res[64][2]) < 1e-5
    assert np.max(np.abs(res[32][1] - res[64][1])) < 1e-6
    assert np.max(np.abs(res[32][0] - res[64][0])) < 1e-7


def test_fp32_256_against_fp64():
    """The 256^3 lattice (the FP32 speed configuration) in both precisions
    for 40 steps: per-field error and conservation at full size."""
    n, steps = 256, 40
    out = {}
    for prec in (64, 32):
        print(f"256^3 fp{prec}, {steps} steps", flush=True)
        b = L.BinaryFluid((n, n, n), tau=1.0, mobility=0.25, precision=prec, fatal=True)
        b.init_spinodal(8361235, 0.0, 0.1)
        s0 = b.conserved_sums()
        b.step(steps)
        rho, u, _ = b.hydro_fields()
        out[prec] = (b.phi_field(), rho, u, s0, b.conserved_sums())
        b.close()
    phi64, rho64, u64 = out[64][:3]
    phi32, rho32, u32 = out[32][:3]
    ep, eu = rel_rms(phi32, phi64), rel_rms(u32, u64)
    print(f"256^3 {steps} steps: phi rel-rms {ep:.2e}, max {np.max(np.abs(phi32 - phi64)):.2e}; "
          f"u rel-rms {eu:.2e}", flush=True)
    assert ep < 1e-6
    assert np.max(np.abs(phi32 - phi64)) < 1e-5
    assert eu < 2e-5
    assert np.max(np.abs(rho32 - rho64)) < 1e-6
    (r0, p0), (r1, p1) = out[32][3], out[32][4]
    assert abs(r1 - r0) / r0 < 1e-9
    assert abs(p1 - p0) / n ** 3 < 1e-8
