"""Diagnostics of the binary-fluid state: structure factor and the domain
length L(t).

L(t) follows the definition the reference's util/length_from_sk.c:88-191
computes (Kendon et al. 2001):

    L = 2 pi  sum_k S(k) dk  /  sum_k k S(k) dk,

with S(k) = |phi_hat(k)|^2 averaged over shells of |k|.  That tool needs
FFTW, which is not available here, and it has two index slips (the
wavenumbers above N/2 are shifted by pi instead of 2 pi, and the mean mode
k = 0 is included in the first shell).  This module evaluates the
definition with the correct periodic wavenumbers k_a = 2 pi m_a / N_a,
m_a in (-N_a/2, N_a/2], shells of width dk = 2 pi / min(N), and the mean
removed.  It is pinned by closed forms (tests/test_analysis.py): a single
cosine of wavelength lam gives L = lam, and a two-mode field gives the
amplitude-weighted ratio.
"""

import math

import numpy as np


def structure_factor(phi, shape):
    """Shell-averaged S(k) of a periodic field.  Returns (k, S, counts)
    with k the shell centres (k > 0 only)."""
    nx, ny, nz = shape
    a = np.asarray(phi, dtype=np.float64).reshape(shape)
    a = a - a.mean()
    F = np.fft.rfftn(a)
    s = (F.real ** 2 + F.imag ** 2)
    kx = 2 * math.pi * np.fft.fftfreq(nx)
    ky = 2 * math.pi * np.fft.fftfreq(ny)
    kz = 2 * math.pi * np.fft.rfftfreq(nz)
    kmod = np.sqrt(kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2)
    # rfftn stores only kz >= 0: interior kz planes stand for +kz and -kz
    weight = np.full(kz.shape, 2.0)
    weight[0] = 1.0
    if nz % 2 == 0:
        weight[-1] = 1.0
    w = np.broadcast_to(weight[None, None, :], s.shape)
    dk = 2 * math.pi / min(shape)
    nbin = int(math.ceil(kmod.max() / dk)) + 1
    idx = np.floor(kmod / dk + 0.5).astype(np.int64)     # shell m: |k| in [(m-1/2)dk, (m+1/2)dk)
    S = np.bincount(idx.ravel(), weights=(s * w).ravel(), minlength=nbin)
    cnt = np.bincount(idx.ravel(), weights=w.ravel(), minlength=nbin)
    k = np.arange(nbin) * dk
    keep = (np.arange(nbin) > 0) & (cnt > 0)
    return k[keep], S[keep] / cnt[keep], cnt[keep]


def domain_length(phi, shape):
    """L = 2 pi sum S(k) / sum k S(k) over shells with k > 0."""
    k, S, _ = structure_factor(phi, shape)
    num = S.sum()
    den = (k * S).sum()
    if den == 0.0:
        return float("inf")
    return 2 * math.pi * num / den
