"""The CPU oracle against golden vectors written by the reference's own build.

The fixtures under tests/golden/ come from oracle/_ref/ref_harness, which
drives the reference's lb_collide (collision.c:143), lb_propagation
(propagation.c:43), field_grad_compute with grad_3d_7pt_fluid_d2
(gradient_3d_7pt_fluid.c:73), phi_force_calculation (phi_force.c:74),
phi_cahn_hilliard (phi_cahn_hilliard.c:213) and field_phi_init_spinodal
(field_phi_init.c:484) on seeded inputs.  The bar is bit-exact equality.
"""

import hashlib

import numpy as np
import pytest

import oracle as O


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def assert_exact(got, want, what):
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, what
    bad = np.flatnonzero(got != want)
    if bad.size:
        i = bad[0]
        rel = np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-300))
        pytest.fail(f"{what}: {bad.size} differ, first at {i}: {got[i]!r} vs {want[i]!r}, max rel {rel:.3e}")


@pytest.mark.parametrize("case,relax,tau", [
    ("collide_m10", O.M10, 0.8), ("collide_bgk", O.BGK, 0.8),
    ("collide_trt", O.TRT, 0.8), ("collide_m10_tau1", O.M10, 1.0)])
def test_collide(golden, case, relax, tau):
    g = golden(case)
    n = int(g["n"])
    p = O.lb_param(tau=tau, relax=relax)
    f, rho, u = O.collide((n, n, n), p, g["f_in"], g["force"])
    assert_exact(f, g["f_out"], "f")
    assert_exact(rho, g["rho"], "rho")
    assert_exact(u, g["u"], "u")


def test_propagate(golden):
    g = golden("prop")
    n = int(g["n"])
    assert_exact(O.propagate((n, n, n), g["f_in"]), g["f_out"], "f")


def _grad_mu(g):
    return tuple(float(v) for v in g["grad_mu"]) if "grad_mu" in g else (0.0, 0.0, 0.0)


@pytest.mark.parametrize("case,method", [("force_stress", 0), ("force_gradmu", 1),
                                         ("force_gradmu_ext", 1)])
def test_gradient_mu_force(golden, case, method):
    """force_gradmu_ext: -phi grad_mu added to the phi grad mu force
    (phi_grad_mu.c:123-160, 509-534)."""
    g = golden(case)
    n = int(g["n"])
    s = (n, n, n)
    fe = O.fe_param(force_method=method, grad_mu=_grad_mu(g))
    grad, delsq = O.grad7(s, g["phi"])
    assert_exact(grad, g["grad"], "grad")
    assert_exact(delsq, g["delsq"], "delsq")
    assert_exact(O.mu(fe, g["phi"], delsq), g["mu"], "mu")
    assert_exact(O.force(s, fe, g["phi"], grad, delsq), g["force"], "force")


@pytest.mark.parametrize("case,order", [("ch_order1", 1), ("ch_order2", 2), ("ch_order3", 3),
                                        ("ch_order4", 4), ("ch_order5", 5), ("ch_gradmu_o2", 2)])
def test_cahn_hilliard(golden, case, order):
    """Orders 1-5 (advection.c:538-1430; order 5 from a reference built with a
    coordinate halo of 3, which its stencil needs) and the external
    chemical-potential gradient on the face fluxes
    (phi_cahn_hilliard.c:1373-1398)."""
    g = golden(case)
    n = int(g["n"])
    s = (n, n, n)
    fe = O.fe_param(order=order, grad_mu=_grad_mu(g))
    _, delsq = O.grad7(s, g["phi_in"])
    assert_exact(O.cahn_hilliard(s, fe, g["phi_in"], delsq, g["u"]), g["phi_out"], "phi")


def test_orders_differ(golden):
    """Each scheme gives its own update of the same input (a case routed to
    the wrong order would fail here as well as above)."""
    outs = [golden(f"ch_order{k}")["phi_out"] for k in range(1, 6)]
    for i in range(5):
        for j in range(i):
            assert np.any(outs[i] != outs[j])


def test_uniform_grad_mu_only_rounds_the_update(golden):
    """A uniform external gradient shifts both faces of a cell alike, so in
    exact arithmetic the divergence is unchanged: against the update without
    it the difference is rounding only."""
    g = golden("ch_gradmu_o2")
    n = int(g["n"])
    s = (n, n, n)
    _, delsq = O.grad7(s, g["phi_in"])
    plain = O.cahn_hilliard(s, O.fe_param(order=2), g["phi_in"], delsq, g["u"])
    d = np.abs(plain - g["phi_out"])
    assert np.any(d != 0.0)
    assert np.max(d) < 64 * np.finfo(float).eps


def _binary(g):
    args = [str(a) for a in g["args"]]
    n, nsteps, seed, order = int(args[1]), int(args[2]), int(args[3]), int(args[4])
    a, b, k, m, tau, phi0, amp = map(float, args[5:12])
    s = (n, n, n)
    lb = O.lb_param(tau=tau)
    stencil = int(g["stencil"]) if "stencil" in g else 7
    conserve = int(g["conserve"]) if "conserve" in g else 0
    method = int(g["force_method"]) if "force_method" in g else 0
    fe = O.fe_param(a=a, b=b, kappa=k, mobility=m, order=order, stencil=stencil,
                    conserve=conserve, force_method=method, grad_mu=_grad_mu(g))
    phi = O.spinodal(s, seed, phi0, amp)
    f = O.equilibrium(s, rho_uniform=1.0)
    return s, nsteps, lb, fe, phi, f


@pytest.mark.parametrize("case,method", [("force27_stress", 0), ("force27_gradmu", 1)])
def test_gradient27_mu_force(golden, case, method):
    """The 27-point gradient (gradient_3d_27pt_fluid.c:219-368) and what
    follows from it, bit-exact against the reference built with it."""
    g = golden(case)
    assert int(g["stencil"]) == 27
    n = int(g["n"])
    s = (n, n, n)
    fe = O.fe_param(force_method=method, stencil=27)
    grad, delsq = O.grad27(s, g["phi"])
    assert_exact(grad, g["grad"], "grad")
    assert_exact(delsq, g["delsq"], "delsq")
    assert_exact(O.mu(fe, g["phi"], delsq), g["mu"], "mu")
    assert_exact(O.force(s, fe, g["phi"], grad, delsq), g["force"], "force")


@pytest.mark.parametrize("case", ["binary16_o1_s20", "binary16_o3_s10", "binary16_o2_s5",
                                  "binary16_g27_s10", "binary16_g27_o3_s10",
                                  "binary16_o4_s10", "binary16_o5_s10",
                                  "binary16_gmu_s10", "binary16_gmu_stress_s10"])
def test_binary_steps_full_fields(golden, case):
    g = golden(case)
    s, nsteps, lb, fe, phi, f = _binary(g)
    if "phi_init" in g:
        assert_exact(phi, g["phi_init"], "phi_init")
    if "phi_10" in g:
        _, p10, _, _ = O.binary_steps(s, 10, lb, fe, f, phi)
        assert_exact(p10, g["phi_10"], "phi at step 10")
    f1, phi1, rho1, u1 = O.binary_steps(s, nsteps, lb, fe, f, phi)
    assert_exact(phi1, g["phi"], "phi")
    assert_exact(rho1, g["rho"], "rho")
    assert_exact(u1, g["u"], "u")
    assert sha(f1) == str(g["f_sha256"])


@pytest.mark.parametrize("case", ["binary16_c1_s20", "binary16_c1_o3_s10", "binary16_c2_s20",
                                  "binary16_c2_o2_s10", "binary16_o4_g27_c1_s10"])
def test_binary_steps_conserve(golden, case):
    """cahn_hilliard_options_conserve 1 (per-site Kahan update with the
    compensation field, phi_cahn_hilliard.c:1181-1218) and 2 (subtraction of
    the drift of the site-order sum from the initial Kahan sum,
    phi_cahn_hilliard.c:1102-1175, 1225-1327; cahn_hilliard_stats.c:59-76),
    bit-exact against the reference run on one thread."""
    g = golden(case)
    s, nsteps, lb, fe, phi, f = _binary(g)
    assert fe.conserve in (1, 2)
    phi_sum = 0.0
    if fe.conserve == 2:
        phi_sum = O.phi_init_sum(phi)
        assert phi_sum == float(g["phi_init_sum"][0])
    f1, phi1, rho1, u1, cs = O.binary_steps_ext(s, nsteps, lb, fe, f, phi, phi_init_sum=phi_sum)
    assert_exact(phi1, g["phi"], "phi")
    assert_exact(rho1, g["rho"], "rho")
    assert_exact(u1, g["u"], "u")
    assert sha(f1) == str(g["f_sha256"])
    if fe.conserve == 1:
        assert_exact(cs, g["csum"], "csum")
        assert np.any(cs != 0.0)
    # the options change the trajectory: the plain update differs
    fe0 = O.fe_param(a=fe.a, b=fe.b, kappa=fe.kappa, mobility=fe.mobility, order=fe.order)
    _, phi_plain, _, _ = O.binary_steps(s, nsteps, lb, fe0, f, phi)
    assert np.any(phi_plain != phi1)


def test_conserve2_pulls_the_sum_to_the_reference_sum():
    """Conserve 2 in closed form: with a reference sum 1e-3 above the
    initial one, one step moves every site by the same (S - S_ref)/N, so
    the site sum lands on the reference sum; the plain step stays put.  A
    planted mistake -- the correction with the wrong sign -- misses."""
    s = (12, 10, 8)
    n = 12 * 10 * 8
    phi = O.spinodal(s, 99, 0.3, 0.4)
    f = O.equilibrium(s, rho_uniform=1.0)
    lb = O.lb_param(tau=0.8)
    target = O.phi_init_sum(phi) + 1.0e-3
    fe2 = O.fe_param(a=-0.0625, b=0.0625, kappa=0.04, mobility=0.25, order=1, conserve=2)
    fe0 = O.fe_param(a=-0.0625, b=0.0625, kappa=0.04, mobility=0.25, order=1)
    _, p2, _, _, _ = O.binary_steps_ext(s, 1, lb, fe2, f, phi, phi_init_sum=target)
    _, p0, _, _ = O.binary_steps(s, 1, lb, fe0, f, phi)
    assert abs(float(np.sum(p2)) - target) < 1e-11
    assert abs(float(np.sum(p0)) - target) > 0.9e-3
    np.testing.assert_allclose(p2 - p0, 1.0e-3 / n, rtol=1e-8)
    wrong = p0 - (p2 - p0)
    assert abs(float(np.sum(wrong)) - target) > 1.9e-3


def test_conserve1_is_the_plain_update_up_to_rounding():
    """Conserve 1 from a zero compensation field: the same update as the
    plain forward step up to rounding (a dropped or sign-flipped flux term
    would differ by the size of a flux, ~1e-3 here), with a compensation
    of at most one ulp of the new phi at every site."""
    s = (8, 6, 6)
    phi = O.spinodal(s, 5, 0.1, 0.5)
    lb = O.lb_param(tau=0.8)
    f = O.equilibrium(s, rho_uniform=1.0)
    fe1 = O.fe_param(a=-0.0625, b=0.0625, kappa=0.04, mobility=0.25, order=2, conserve=1)
    fe0 = O.fe_param(a=-0.0625, b=0.0625, kappa=0.04, mobility=0.25, order=2)
    _, p1, _, _, cs = O.binary_steps_ext(s, 1, lb, fe1, f, phi)
    _, p0, _, _ = O.binary_steps(s, 1, lb, fe0, f, phi)
    eps = np.finfo(float).eps
    assert np.max(np.abs(p1 - p0)) <= 8 * eps * np.max(np.abs(phi))
    assert np.max(np.abs(p1 - phi)) > 1e-4
    assert np.any(cs != 0.0)
    assert np.all(np.abs(cs) <= np.spacing(np.abs(p1)))


@pytest.mark.parametrize("case", ["binary32_o1_s100", "binary32_o3_s200", "binary32_m045_s20"])
def test_binary32_long(golden, case):
    g = golden(case)
    s, nsteps, lb, fe, phi, f = _binary(g)
    f1, phi1, rho1, u1 = O.binary_steps(s, nsteps, lb, fe, f, phi)
    for key, a in (("phi", phi1), ("rho", rho1), ("u", u1), ("f", f1)):
        assert sha(a) == str(g[key + "_sha256"]), key


def test_stated_c2_mobility_is_unstable_in_the_reference(golden):
    """At M = 0.45 (with A = -0.0625, B = 0.0625, kappa = 0.04) the explicit
    Cahn-Hilliard step amplifies the grid-scale mode: the reference's own run
    is non-finite by step 40, and so is the oracle's."""
    g = golden("binary32_m045_s40")
    assert not bool(g["phi_finite"])
    s, nsteps, lb, fe, phi, f = _binary(g)
    _, phi1, _, _ = O.binary_steps(s, nsteps, lb, fe, f, phi)
    assert not np.isfinite(phi1).all()


def test_single_fluid_c1(golden):
    """C1: plain D3Q19 MRT 32^3, tau 0.8, 100 steps from a random-velocity
    equilibrium; bit-exact f, rho, u after 100 steps."""
    g = golden("single32_s100")
    args = [str(a) for a in g["args"]]
    n, nsteps, seed, tau = int(args[1]), int(args[2]), int(args[3]), float(args[4])
    s = (n, n, n)
    u0 = O.random_velocity(s, seed)
    assert sha(u0) == str(g["u_init_sha256"])
    f0 = O.equilibrium(s, u=u0, rho_uniform=1.0)
    assert sha(f0) == str(g["f_init_sha256"])
    f, rho, u = O.single_steps(s, nsteps, O.lb_param(tau=tau), f0)
    for key, a in (("f", f), ("rho", rho), ("u", u)):
        assert sha(a) == str(g[key + "_sha256"]), key
