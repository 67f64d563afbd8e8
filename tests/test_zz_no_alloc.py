"""ldw_nn_restore_sums_noalloc, the entry point that allocates nothing
(its partial sums live in a fixed device array), against
ldw_nn_restore_sums on the same inputs, through the C ABI: bit for bit, on
two calls in a row, the second with a smaller input, so a stale partial of
the first call would show.  Sizes beyond the fixed scratch are refused."""

import ctypes

import numpy as np
import pytest

import ludwig_amd as L

pytestmark = pytest.mark.gpu


def _vp(a):
    return ctypes.c_void_p(a.ptr)


def _case(lib, rng, ns, nplane):
    now = rng.standard_normal(nplane * ns).astype(np.float32)
    out = (now + 0.02 + 0.1 * rng.standard_normal(nplane * ns)).astype(np.float32)
    dnow = L.DeviceArray.from_host(now, np.float32)
    d_a = L.DeviceArray.from_host(out, np.float32)
    d_b = L.DeviceArray.from_host(out, np.float32)
    assert lib.ldw_nn_restore_sums(_vp(dnow), _vp(d_a), ns, nplane) == 0
    assert lib.ldw_nn_restore_sums_noalloc(_vp(dnow), _vp(d_b), ns, nplane) == 0
    a, b = d_a.download(), d_b.download()
    for d in (dnow, d_a, d_b):
        d.free()
    return now, out, a, b


def test_noalloc_restore_sums_equals_the_allocating_entry_point():
    lib = L.lib()
    rng = np.random.default_rng(11)
    # first call: 64 planes (the maximum) of 70001 sites; second: 3 planes
    # of 1000 sites, so most partial slots keep the first call's values
    for ns, nplane in ((70001, 64), (1000, 3)):
        print(f"noalloc restore sums: ns={ns} nplane={nplane}", flush=True)
        now, out, a, b = _case(lib, rng, ns, nplane)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (ns, nplane)
        # and the result is the restore: each plane's sum back to that of `now`
        got = b.reshape(nplane, ns).astype(np.float64)
        want = now.reshape(nplane, ns).astype(np.float64).sum(1)
        assert np.allclose(got.sum(1), want, atol=ns * 2e-7)


def test_noalloc_restore_sums_refuses_more_planes_than_its_scratch():
    lib = L.lib()
    L.lib().ldw_error_mode_set(0)
    try:
        d = L.DeviceArray.from_host(np.zeros(65 * 8, np.float32), np.float32)
        assert lib.ldw_nn_restore_sums_noalloc(_vp(d), _vp(d), 8, 65) == -1
        assert b"at most 64" in lib.ldw_last_error()
        assert lib.ldw_nn_restore_sums_noalloc(_vp(d), _vp(d), 0, 1) == -1
        d.free()
    finally:
        lib.ldw_error_mode_set(1)
