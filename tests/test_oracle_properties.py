"""Properties that pin the CPU oracle independently of the golden vectors:
the reference's own known answers (tests/unit/test_noise.c, the committed
regression log serial-spin-fd1.log) and what the mathematics fixes
(closed forms, invariants, textbook routines).  Each check is chosen so that
a plausible slip -- a dropped term, a wrong sign, a transposed index --
fails it."""

import math

import numpy as np
import pytest

import oracle as O

EPS = np.finfo(np.float64).eps


# --- lattice generator: tests/unit/test_noise.c:96-123 ----------------------

def test_ns_uniform_known_answers():
    r, s = O.ns_uniform([0, 0, 0, 0])
    assert s == [1234567, 0, 0, 0] and r == 1234567
    r, s = O.ns_uniform([123, 456, 78, 9])
    assert s == [9730054, 1905505352, 2883582, 162000]
    assert r == 1915204894


def test_spinodal_matches_reference_log_line():
    """tests/regression/d3q19-short/serial-spin-fd1.log, step 0:
    [phi]  3.1484764e+00  1.2010484e-05 8.3289934e-04 -4.9999916e-02 4.9999705e-02
    (64^3, seed 8361235, phi0 0, noise 0.1)."""
    phi = O.spinodal((64, 64, 64), 8361235, 0.0, 0.1)
    total = O.kahan_sum(phi)
    mean = total / phi.size
    var = np.sum(phi * phi) / phi.size - mean * mean
    assert f"{total:14.7e}" == " 3.1484764e+00"
    assert f"{mean:14.7e}" == " 1.2010484e-05"
    assert f"{var:14.7e}" == " 8.3289934e-04"
    assert f"{phi.min():14.7e}" == "-4.9999916e-02"
    assert f"{phi.max():14.7e}" == " 4.9999705e-02"


# --- D3Q19 model: tests/unit/test_lb_model.c:103-377 ------------------------

def test_d3q19_tables():
    cv, wv, ma, mi = O.tables()
    assert cv.shape == (19, 3)
    for p in range(1, 19):
        assert (cv[19 - p] == -cv[p]).all()            # opposite of p >= 1 is 19 - p
    assert abs(wv.sum() - 1.0) < 5 * EPS
    assert np.allclose(wv @ cv, 0.0, atol=5 * EPS)
    cc = np.einsum("p,pa,pb->ab", wv, cv, cv)
    assert np.allclose(cc, np.eye(3) / 3.0, atol=5 * EPS)        # cs^2 = 1/3
    # rows are orthogonal under the weights; mi is the inverse of ma
    gram = ma @ np.diag(wv) @ ma.T
    assert np.allclose(gram - np.diag(np.diag(gram)), 0.0, atol=5 * EPS)
    assert np.allclose(mi @ ma, np.eye(19), atol=5 * EPS)
    assert np.allclose(ma @ mi, np.eye(19), atol=5 * EPS)
    # hydrodynamic rows: 1, c, cc - 1/3
    assert (ma[0] == 1).all()
    assert (ma[1:4] == cv.T).all()
    assert np.allclose(ma[4], cv[:, 0] ** 2 - 1.0 / 3.0, atol=EPS)
    assert (ma[5] == cv[:, 0] * cv[:, 1]).all()


# --- collision invariants and closed forms -----------------------------------

def _random_f(shape, seed, amp=1e-3):
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    rho = 1.0 + 0.05 * rng.standard_normal(n)
    u = 0.02 * rng.standard_normal(3 * n)
    f = O.equilibrium(shape, rho=rho, u=u)
    return f + amp * rng.standard_normal(f.size)


@pytest.mark.parametrize("relax", [O.M10, O.BGK, O.TRT])
def test_collision_conserves_mass_and_adds_force(relax):
    shape = (4, 5, 6)
    n = 4 * 5 * 6
    f = _random_f(shape, 1)
    F = 1e-4 * np.random.default_rng(2).standard_normal(3 * n)
    cv, _, _, _ = O.tables()
    f1, rho, u = O.collide(shape, O.lb_param(tau=0.8, relax=relax), f, F)
    fr, f1r = f.reshape(19, n), f1.reshape(19, n)
    assert np.allclose(f1r.sum(0), fr.sum(0), rtol=0, atol=20 * EPS)
    j0 = cv.T.astype(float) @ fr
    j1 = cv.T.astype(float) @ f1r
    assert np.allclose(j1, j0 + F.reshape(3, n), rtol=0, atol=20 * EPS)
    # u = (j + F/2)/rho (collision.c:370-375)
    assert np.allclose(u.reshape(3, n), (j0 + 0.5 * F.reshape(3, n)) / fr.sum(0), rtol=1e-14)
    assert np.allclose(rho, fr.sum(0), rtol=0, atol=20 * EPS)


def test_equilibrium_is_a_fixed_point_without_force():
    shape = (3, 4, 5)
    n = 60
    rng = np.random.default_rng(3)
    rho = 1.0 + 0.1 * rng.random(n)
    u = 0.03 * (rng.random(3 * n) - 0.5)
    feq = O.equilibrium(shape, rho=rho, u=u)
    f1, _, _ = O.collide(shape, O.lb_param(tau=0.7), feq)
    assert np.allclose(f1, feq, rtol=0, atol=1e-15)


def test_shear_mode_relaxes_by_one_minus_omega():
    """A pure s_xy perturbation on a rest equilibrium decays by (1 - omega)
    per collision, omega = 1/tau (M10: ghosts to zero)."""
    shape = (2, 2, 2)
    n = 8
    _, _, ma, mi = O.tables()
    feq = O.equilibrium(shape, rho_uniform=1.0).reshape(19, n)
    delta = 1e-4
    f = feq + delta * mi[:, 5][:, None]            # add shear mode 5 (s_xy)
    tau = 0.9
    f1, _, _ = O.collide(shape, O.lb_param(tau=tau), f.ravel())
    m1 = ma @ f1.reshape(19, n)
    assert np.allclose(m1[5], delta * (1.0 - 1.0 / tau), rtol=1e-10, atol=1e-18)


def test_propagation_is_the_periodic_shift():
    shape = (3, 4, 5)
    n = 60
    cv, _, _, _ = O.tables()
    f = np.arange(19 * n, dtype=np.float64)
    g = O.propagate(shape, f).reshape(19, *shape)
    fr = f.reshape(19, *shape)
    for p in range(19):
        assert (g[p] == np.roll(fr[p], shift=tuple(cv[p]), axis=(0, 1, 2))).all()


# --- gradient, mu, stress, force ---------------------------------------------

def test_gradient_closed_forms():
    n = 16
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    k = 2 * math.pi * 3 / n
    phi = np.cos(k * X) + 0.5 * np.sin(2 * k * Y) - 0.25 * np.cos(k * Z)
    g, d = O.grad7((n, n, n), phi.ravel())
    g = g.reshape(3, n, n, n)
    assert np.allclose(g[0], -np.sin(k) * np.sin(k * X), atol=1e-14)
    assert np.allclose(g[1], 0.5 * np.sin(2 * k) * np.cos(2 * k * Y), atol=1e-14)
    lap = (2 * (np.cos(k) - 1) * np.cos(k * X) + 0.5 * 2 * (np.cos(2 * k) - 1) * np.sin(2 * k * Y)
           - 0.25 * 2 * (np.cos(k) - 1) * np.cos(k * Z))
    assert np.allclose(d.reshape(n, n, n), lap, atol=1e-14)


def test_mu_and_stress_closed_forms():
    fe = O.fe_param(a=-0.1, b=0.2, kappa=0.03)
    rng = np.random.default_rng(5)
    n = 50
    phi, delsq = rng.standard_normal(n), rng.standard_normal(n)
    grad = rng.standard_normal(3 * n)
    mu = O.mu(fe, phi, delsq)
    assert np.allclose(mu, -0.1 * phi + 0.2 * phi ** 3 - 0.03 * delsq, rtol=1e-14)
    P = O.stress(fe, phi, grad, delsq).reshape(3, 3, n)
    g = grad.reshape(3, n)
    p0 = 0.5 * -0.1 * phi ** 2 + 0.75 * 0.2 * phi ** 4 - 0.03 * phi * delsq - 0.5 * 0.03 * (g * g).sum(0)
    want = p0[None, None, :] * np.eye(3)[:, :, None] + 0.03 * g[:, None, :] * g[None, :, :]
    assert np.allclose(P, want, rtol=1e-13, atol=1e-15)


def test_total_stress_divergence_force_vanishes():
    """F = -div P telescopes to zero on a periodic box."""
    shape = (12, 10, 8)
    phi = 0.8 * (np.random.default_rng(6).random(960) - 0.5)
    fe = O.fe_param(force_method=0)
    g, d = O.grad7(shape, phi)
    F = O.force(shape, fe, phi, g, d).reshape(3, -1)
    assert np.allclose(F.sum(1), 0.0, atol=1e-13)
    assert np.abs(F).max() > 1e-3


def test_phi_grad_mu_force_closed_form():
    """F_a = -phi (mu(x + e_a) - mu(x - e_a)) / 2 with mu from the 7-point
    Laplacian."""
    shape = (12, 10, 8)
    phi = 0.8 * (np.random.default_rng(6).random(960) - 0.5)
    fe = O.fe_param(force_method=1)
    g, d = O.grad7(shape, phi)
    F = O.force(shape, fe, phi, g, d).reshape(3, *shape)
    mu = O.mu(fe, phi, d).reshape(shape)
    p = phi.reshape(shape)
    for a in range(3):
        want = -p * 0.5 * (np.roll(mu, -1, a) - np.roll(mu, 1, a))
        assert np.allclose(F[a], want, rtol=1e-13, atol=1e-16)


# --- Cahn-Hilliard ------------------------------------------------------------

@pytest.mark.parametrize("order", [1, 2, 3])
def test_cahn_hilliard_conserves_phi(order):
    shape = (10, 12, 8)
    rng = np.random.default_rng(7)
    phi = 0.6 * (rng.random(960) - 0.5)
    u = 0.02 * (rng.random(3 * 960) - 0.5)
    fe = O.fe_param(order=order, mobility=0.2)
    _, d = O.grad7(shape, phi)
    phi1 = O.cahn_hilliard(shape, fe, phi, d, u)
    assert abs(math.fsum(phi1) - math.fsum(phi)) < 1e-13


def test_cahn_hilliard_linear_dispersion():
    """At u = 0 and small amplitude, a Fourier mode obeys
    phi' = phi (1 - M khat^2 (A + kappa khat^2)), khat^2 = sum 2 (1 - cos k)."""
    n = 16
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    kx, ky = 2 * math.pi * 2 / n, 2 * math.pi * 1 / n
    eps = 1e-7
    phi = eps * np.cos(kx * X + ky * Y)
    a, kappa, m = -0.0625, 0.04, 0.25
    fe = O.fe_param(a=a, b=0.0625, kappa=kappa, mobility=m)
    _, d = O.grad7((n, n, n), phi.ravel())
    phi1 = O.cahn_hilliard((n, n, n), fe, phi.ravel(), d, np.zeros(3 * n ** 3))
    k2 = 2 * (1 - math.cos(kx)) + 2 * (1 - math.cos(ky))
    growth = 1 - m * k2 * (a + kappa * k2)
    assert np.allclose(phi1.reshape(n, n, n), growth * phi, rtol=0, atol=1e-12 * eps)


def test_upwind_shift_with_uniform_velocity():
    """First-order upwind with u_x = 1 (no diffusion, M = 0) moves phi one
    site per step in x exactly."""
    shape = (8, 4, 4)
    phi = np.random.default_rng(8).random(128)
    u = np.zeros(3 * 128)
    u[:128] = 1.0
    fe = O.fe_param(order=1, mobility=0.0)
    _, d = O.grad7(shape, phi)
    phi1 = O.cahn_hilliard(shape, fe, phi, d, u)
    assert np.allclose(phi1.reshape(shape), np.roll(phi.reshape(shape), 1, axis=0), atol=1e-15)


# --- compensated sums: tests/unit/test_util_sum.c ------------------------------

def test_compensated_sums():
    a = np.array([1.0, 1e100, 1.0, -1e100] * 1000)
    assert math.fsum(a) == 2000.0
    assert O.klein_sum(a) == 2000.0
    rng = np.random.default_rng(9)
    b = rng.standard_normal(100000) * 10.0 ** rng.integers(-8, 8, 100000)
    exact = math.fsum(b)
    assert abs(O.kahan_sum(b) - exact) <= 4 * EPS * np.abs(b).sum() * 1e-6 + abs(exact) * 4 * EPS
    assert abs(O.klein_sum(b) - exact) <= abs(exact) * 2 * EPS


def test_gradient27_closed_form_on_a_plane_wave():
    """27-point stencil on phi = cos(k.r): the 26 neighbours sum to
    prod_a (1 + 2 cos k_a) - 1 times phi, so
    del2 phi = (prod_a (1 + 2 cos k_a) - 27)/9 phi and
    d_x phi = -(1/9) sin k_x (1 + 2 cos k_y)(1 + 2 cos k_z) sin(k.r)."""
    nx, ny, nz = 10, 12, 14
    m = (3, -2, 5)
    k = [2 * np.pi * m[0] / nx, 2 * np.pi * m[1] / ny, 2 * np.pi * m[2] / nz]
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    arg = k[0] * X + k[1] * Y + k[2] * Z
    phi = np.cos(arg).ravel()
    grad, delsq = O.grad27((nx, ny, nz), phi)
    c = [1 + 2 * np.cos(ka) for ka in k]
    assert np.allclose(delsq, (c[0] * c[1] * c[2] - 27) / 9 * phi, atol=1e-13)
    s = np.sin(arg).ravel()
    n = phi.size
    assert np.allclose(grad[:n], -np.sin(k[0]) * c[1] * c[2] / 9 * s, atol=1e-13)
    assert np.allclose(grad[n:2 * n], -np.sin(k[1]) * c[0] * c[2] / 9 * s, atol=1e-13)
    assert np.allclose(grad[2 * n:], -np.sin(k[2]) * c[0] * c[1] / 9 * s, atol=1e-13)
