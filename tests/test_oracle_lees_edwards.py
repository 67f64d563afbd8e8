"""The CPU oracle's Lees-Edwards plane code (oracle/le_oracle.c), pinned.

Against the golden vectors of the reference's own build with planes
(tests/golden/le_*.npz, written by oracle/ref_harness.c): the force with
the 7- and 27-point gradients, one Cahn-Hilliard update for advection
orders 1-3, the LB plane conditions, and whole step loops, all bit-exact.

Against what the mathematics fixes, for what the vectors leave open:
  - planes that do not move (uy = 0) leave the Cahn-Hilliard update and the
    LB distributions exactly as without planes, and the flux-form force
    equal to the stress-divergence force to rounding;
  - the LB plane conditions conserve mass, and with an integer
    displacement move whole rows;
  - the shear-profile start has density rho0 and the piecewise linear
    velocity profile whose jumps equal the plane speed;
  - plane geometries the reference refuses are refused.
A dropped term, a wrong sign, a plane off by one site or a displacement
in the wrong direction breaks one of these.
"""

import hashlib

import numpy as np
import pytest

import oracle as O


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def assert_exact(got, want, what):
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape, what
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{what}: {bad.size}/{got.size} differ, first {bad[0]}: {got[bad[0]]!r} vs {want[bad[0]]!r}"


def _le(g, nhalo=2):
    nt0 = int(g["le_nt0"]) if "le_nt0" in g else 0
    return O.LeParam(int(g["le_planes"]), nhalo, float(g["le_uy"]), nt0)


@pytest.mark.parametrize("case,stencil", [("le_force16", 7), ("le_force16_g27", 27)])
def test_le_force_golden(golden, case, stencil):
    g = golden(case)
    n = int(g["n"])
    a, b, k = map(float, [str(x) for x in g["args"]][4:7])
    fe = O.fe_param(a=a, b=b, kappa=k, stencil=stencil)
    t = float(int(g["le_t0"]) - 1)            # physics_control_time after T0 steps
    assert_exact(O.le_force((n, n, n), fe, _le(g), t, g["phi"]), g["force"], "force")


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_le_cahn_hilliard_golden(golden, order):
    g = golden(f"le_ch16_o{order}")
    n = int(g["n"])
    a, b, k, m = map(float, [str(x) for x in g["args"]][4:8])
    fe = O.fe_param(a=a, b=b, kappa=k, mobility=m, order=order)
    t = float(int(g["le_t0"]) - 1)
    phi, _ = O.le_cahn_hilliard((n, n, n), fe, _le(g), t, g["phi_in"], g["u"])
    assert_exact(phi, g["phi_out"], "phi")


def test_le_lb_bc_golden(golden):
    g = golden("le_lbbc16")
    n = int(g["n"])
    f = O.le_lb_bc((n, n, n), _le(g), int(g["le_t0"]), g["f_in"])
    assert_exact(f, g["f_out"], "f")


@pytest.mark.parametrize("case", ["le_binary16_o1_s20", "le_binary16_o3_g27_s10",
                                  "le_binary24_p3_o2_s12", "le_binary16_c1_s10",
                                  "le_binary16_t0_s12", "le_binary16_o4_s10",
                                  "le_binary16_gmu_s10"])
def test_le_binary_steps_golden(golden, case):
    g = golden(case)
    args = [str(a) for a in g["args"]]
    n, nsteps, seed, order = int(args[1]), int(args[2]), int(args[3]), int(args[4])
    a, b, k, m, tau, phi0, amp = map(float, args[5:12])
    s = (n, n, n)
    lb = O.lb_param(tau=tau)
    gm = tuple(float(v) for v in g["grad_mu"]) if "grad_mu" in g else (0.0, 0.0, 0.0)
    fe = O.fe_param(a=a, b=b, kappa=k, mobility=m, order=order, stencil=int(g["stencil"]),
                    conserve=int(g["conserve"]), grad_mu=gm)
    le = _le(g)
    phi = O.spinodal(s, seed, phi0, amp)
    if int(g["le_profile"]):
        f = O.le_shear_profile(s, le, 1.0, (tau - 0.5) / 3.0)
    else:
        f = O.equilibrium(s, rho_uniform=1.0)
    f1, phi1, rho1, u1, cs = O.le_binary_steps(s, nsteps, lb, fe, le, f, phi)
    assert_exact(phi1, g["phi"], "phi")
    assert_exact(rho1, g["rho"], "rho")
    assert_exact(u1, g["u"], "u")
    assert sha(f1) == str(g["f_sha256"]), "f"
    if "csum" in g:
        assert_exact(cs, g["csum"], "csum")


def _random(shape, seed, n):
    rng = np.random.default_rng(seed)
    return rng.uniform(-0.8, 0.8, n * int(np.prod(shape)))


def test_planes_at_rest_change_nothing():
    """uy = 0: the Cahn-Hilliard update equals the plane-free one bit for
    bit (the averaged plane faces are equal on both sides), and the force
    equals the stress-divergence force to rounding (the flux form sums the
    same face averages in another order)."""
    s = (24, 10, 8)
    le = O.LeParam(3, 2, 0.0, 0)
    phi = _random(s, 1, 1)
    u = 0.02 * _random(s, 2, 3)
    for order in (1, 2, 3):
        fe = O.fe_param(order=order, mobility=0.2)
        got, _ = O.le_cahn_hilliard(s, fe, le, 5.0, phi, u)
        _, delsq = O.grad7(s, phi)
        assert_exact(got, O.cahn_hilliard(s, fe, phi, delsq, u), f"order {order}")
    fe = O.fe_param()
    grad, delsq = O.grad7(s, phi)
    want = O.force(s, fe, phi, grad, delsq)
    got = O.le_force(s, fe, le, 5.0, phi)
    assert np.abs(got - want).max() <= 1e-14 * np.abs(want).max()
    f = np.abs(_random(s, 3, 19)) + 0.01
    assert_exact(O.le_lb_bc(s, le, 7, f), f, "LB planes at rest")


def test_moving_planes_change_only_the_planes():
    """The force and phi change only within reach of the planes: 2 sites
    for the gradient, 1 more for the stress divergence."""
    s = (32, 12, 6)
    le = O.LeParam(2, 2, 0.37, 0)
    rest = O.LeParam(2, 2, 0.0, 0)
    phi = _random(s, 4, 1)
    fe = O.fe_param()
    d = (O.le_force(s, fe, le, 8.0, phi) != O.le_force(s, fe, rest, 8.0, phi)).reshape(3, 32, -1)
    xs = set(np.flatnonzero(d.any(axis=(0, 2))))
    near = {x for L in (8, 24) for x in range(L - 3, L + 3)}     # 0-based x of L-2 .. L+3
    assert xs and xs <= near, sorted(xs)


def test_lb_planes_conserve_mass_and_shift_rows():
    """Mass is conserved for any displacement.  With uy = 0.5 the plane has
    slid exactly 2 rows at time step 4 and none at step 0, and the
    reprojection does not depend on time: the crossing populations at step 4
    are those of step 0 taken 2 rows further along y below the plane
    (cx = +1) and 2 rows back above it (cx = -1)."""
    s = (16, 12, 5)
    f = np.abs(_random(s, 5, 19)) + 0.02
    for uy, ts in ((0.37, 9), (0.13, 40), (0.5, 4)):
        out = O.le_lb_bc(s, O.LeParam(2, 2, uy, 0), ts, f)
        assert abs(out.sum() - f.sum()) <= 1e-12 * f.sum()
    le = O.LeParam(2, 2, 0.5, 0)
    a4 = O.le_lb_bc(s, le, 4, f).reshape(19, 16, 12, 5)
    a0 = O.le_lb_bc(s, le, 0, f).reshape(19, 16, 12, 5)
    f4 = f.reshape(19, 16, 12, 5)
    for L in (4, 12):                        # planes between x = L and L + 1 (1-based)
        for p in (1, 2, 3, 4, 5):            # cx = +1, at x = L
            assert np.array_equal(a4[p, L - 1], np.roll(a0[p, L - 1], -2, axis=0))
            assert not np.array_equal(a0[p, L - 1], f4[p, L - 1])
        for p in (14, 15, 16, 17, 18):       # cx = -1, at x = L + 1
            assert np.array_equal(a4[p, L], np.roll(a0[p, L], +2, axis=0))
    untouched = [x for x in range(16) if x not in (3, 4, 11, 12)]
    assert np.array_equal(a4[:, untouched], f4[:, untouched])


def test_shear_profile_moments():
    nx, ny, nz, uy, rho0, eta = 32, 4, 3, 0.05, 1.0, 0.1
    f = O.le_shear_profile((nx, ny, nz), O.LeParam(2, 2, uy, 0), rho0, eta).reshape(19, nx, ny, nz)
    # D3Q19 velocities (lb_d3q19.h:24-32)
    cv = np.array([[0, 0, 0], [1, 1, 0], [1, 0, 1], [1, 0, 0], [1, 0, -1], [1, -1, 0],
                   [0, 1, 1], [0, 1, 0], [0, 1, -1], [0, 0, 1], [0, 0, -1], [0, -1, 1],
                   [0, -1, 0], [0, -1, -1], [-1, 1, 0], [-1, 0, 1], [-1, 0, 0], [-1, 0, -1],
                   [-1, -1, 0]])
    rho = f.sum(axis=0)
    mom = np.einsum("pa,pxyz->axyz", cv, f)
    assert np.abs(rho - rho0).max() < 1e-15
    x = np.arange(1, nx + 1) - 0.5
    want = x * uy * 2 / nx - uy * np.floor((nx / 4 + x) / (nx / 2))
    assert np.abs(mom[1] - want[:, None, None]).max() < 1e-15
    assert np.abs(mom[0]).max() < 1e-16 and np.abs(mom[2]).max() < 1e-16
    jumps = np.diff(mom[1, :, 0, 0])
    big = np.flatnonzero(np.abs(jumps) > 0.01)
    # across each plane (between x = 8 and 9, 24 and 25) the slope continues
    # and the plane speed is taken off
    assert list(big) == [7, 23] and np.allclose(jumps[big], uy * 2 / nx - uy)


def test_refused_geometries():
    fe = O.fe_param()
    phi = np.zeros(16 ** 3)
    with pytest.raises(ValueError):
        O.le_force((16, 16, 16), fe, O.LeParam(3, 2, 0.1, 0), 0.0, phi)    # 3 does not divide 16
    with pytest.raises(ValueError):
        O.le_force((16, 16, 16), fe, O.LeParam(8, 2, 0.1, 0), 0.0, phi)    # plane at x = 1
