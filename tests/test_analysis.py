"""L(t) of ludwig_amd.analysis against closed forms.  The reference's
length_from_sk.c cannot be built here (it needs FFTW), so this part is pinned
by the mathematics only: a single Fourier mode of wavelength lam lies in one
shell, giving L = lam exactly; a two-mode field gives the amplitude-weighted
ratio; L is translation invariant and scales with the pattern."""

import math

import numpy as np
import pytest

from ludwig_amd.analysis import domain_length, structure_factor


def grid(n):
    x = np.arange(n)
    return np.meshgrid(x, x, x, indexing="ij")


@pytest.mark.parametrize("n,m", [(32, 4), (64, 8), (48, 3)])
def test_single_mode_gives_its_wavelength(n, m):
    X, Y, Z = grid(n)
    phi = np.cos(2 * math.pi * m * X / n) + 0.3      # mean is removed
    assert abs(domain_length(phi.ravel(), (n, n, n)) - n / m) < 1e-9


def test_two_modes_closed_form():
    n = 64
    X, Y, Z = grid(n)
    a1, m1, a2, m2 = 1.0, 4, 0.5, 10
    phi = a1 * np.cos(2 * math.pi * m1 * X / n) + a2 * np.cos(2 * math.pi * m2 * Y / n)
    k1, k2 = 2 * math.pi * m1 / n, 2 * math.pi * m2 / n
    # both modes sit on their shell centres; S ~ amplitude^2 / shell count
    k, S, cnt = structure_factor(phi.ravel(), (n, n, n))
    i1 = np.argmin(abs(k - k1))
    i2 = np.argmin(abs(k - k2))
    s1, s2 = S[i1], S[i2]
    want = 2 * math.pi * (s1 + s2) / (k1 * s1 + k2 * s2)
    assert abs(domain_length(phi.ravel(), (n, n, n)) - want) < 1e-9
    assert s1 / s2 == pytest.approx((a1 / a2) ** 2 * cnt[i2] / cnt[i1])


def test_translation_invariance_and_scaling():
    n = 32
    rng = np.random.default_rng(1)
    phi = rng.standard_normal((n, n, n))
    L0 = domain_length(phi.ravel(), (n, n, n))
    L1 = domain_length(np.roll(phi, (3, 5, 7), axis=(0, 1, 2)).ravel(), (n, n, n))
    assert abs(L0 - L1) < 1e-9 * L0
    # a pattern repeated on a lattice twice as fine (each value duplicated
    # along every axis) has twice the length
    big = np.repeat(np.repeat(np.repeat(phi, 2, 0), 2, 1), 2, 2)
    assert domain_length(big.ravel(), (2 * n,) * 3) > 1.8 * L0
