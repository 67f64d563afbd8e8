"""Replica ensemble (ldw_ensemble_*) through the C ABI.

The bar: every replica is bit-identical to a single lattice of the same
size, parameters and seed advanced by ldw_step_binary (FP32-shifted and
FP64), for every tile shape of the persistent kernel, with and without the
HIP graph; FP64 replicas also match the CPU oracle bit for bit.
"""

import numpy as np
import pytest

import ludwig_amd as L
from ludwig_amd.ensemble import Ensemble, draw_replicas, replica
import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(64, 4), (32, 8), (16, 16), (64, 8), (128, 4)]


def assert_exact(got, want, what):
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape, what
    bad = np.flatnonzero(got != want)
    if bad.size:
        i = bad[0]
        pytest.fail(f"{what}: {bad.size}/{got.size} differ, first {i}: {got[i]!r} vs {want[i]!r}")


def single_run(spec, nsteps, precision, order=1, force_method=L.FORCE_STRESS_DIVERGENCE):
    shape = (spec.nx, spec.ny, spec.nz)
    b = L.BinaryFluid(shape, a=spec.a, b=spec.b, kappa=spec.kappa, mobility=spec.mobility,
                      order=order, force_method=force_method, eta=spec.eta_shear,
                      eta_bulk=spec.eta_bulk, rho0=spec.rho0, precision=precision)
    b.init_spinodal(spec.seed, spec.phi0, spec.noise, rho0=spec.rho0)
    b.step(nsteps)
    phi = b.phi_field()
    rho, u, _ = b.hydro_fields()
    f = b.f()
    b.close()
    return phi, rho, u, f


def small_specs():
    return [replica((16, 16, 16), 101, phi0=0.05, kappa=0.03, mobility=0.2),
            replica((20, 12, 24), 202, phi0=-0.1, kappa=0.05, mobility=0.15),
            replica((9, 17, 33), 303, phi0=0.0, kappa=0.04, mobility=0.25)]


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("tile", SHAPES)
def test_replicas_equal_single_lattice_runs(precision, tile):
    specs = small_specs()
    nsteps = 12
    print(f"ensemble tile {tile} fp{precision} sizes {[(s.nx, s.ny, s.nz) for s in specs]}", flush=True)
    e = Ensemble(specs, precision=precision)
    e.set_config(tile[0], tile[1], xplanes=5)
    e.step(nsteps)
    assert e.time == nsteps
    for r, s in enumerate(specs):
        phi, rho, u, f = single_run(s, nsteps, precision)
        ephi, erho, eu = e.snapshot(r)
        assert_exact(ephi, phi, f"replica {r} phi")
        assert_exact(eu, u, f"replica {r} u")
        assert_exact(erho, rho, f"replica {r} rho")
        assert_exact(e.f(r), f, f"replica {r} f")
    e.close()


def test_fp64_replicas_equal_oracle():
    specs = small_specs()[:2]
    nsteps = 6
    e = Ensemble(specs, precision=64, order=3)
    e.set_config(16, 16, xplanes=7)
    e.step(nsteps)
    for r, s in enumerate(specs):
        shape = (s.nx, s.ny, s.nz)
        f0 = O.equilibrium(shape, rho_uniform=1.0)
        phi0 = O.spinodal(shape, s.seed, s.phi0, s.noise)
        f, phi, rho, u = O.binary_steps(shape, nsteps, O.lb_param(eta=s.eta_shear),
                                        O.fe_param(a=s.a, b=s.b, kappa=s.kappa, mobility=s.mobility,
                                                   order=3), f0, phi0)
        ephi, erho, eu = e.snapshot(r)
        assert_exact(ephi, phi, f"replica {r} phi")
        assert_exact(eu, u, f"replica {r} u")
        assert_exact(e.f(r), f, f"replica {r} f")
    e.close()


def test_graph_direct_tune_and_split_steps_agree():
    specs = small_specs()
    ref = Ensemble(specs, precision=32)
    ref.set_config(32, 8, xplanes=4)
    ref.use_graph(False)
    ref.step(9)
    e = Ensemble(specs, precision=32)
    rows = e.tune(reps=2)            # tuning must not advance the state
    assert rows.shape[0] >= 3 and np.all(rows[:, 4] > 0)
    for r in range(len(specs)):
        assert_exact(e.snapshot(r)[0], O.spinodal((specs[r].nx, specs[r].ny, specs[r].nz),
                                                  specs[r].seed, specs[r].phi0, specs[r].noise)
                     .astype(np.float32).astype(np.float64), f"replica {r} phi after tune")
    e.step(4)
    e.step(5)                        # odd split: both graphs (parities) used
    assert e.time == 9
    for r in range(len(specs)):
        a, b = e.snapshot(r), ref.snapshot(r)
        for k, name in enumerate(("phi", "rho", "u")):
            assert_exact(a[k], b[k], f"replica {r} {name}")
    cfg = e.config()
    assert cfg["tile"] in SHAPES and cfg["tuned_ms"] > 0
    e.close()
    ref.close()


def test_conservation_and_device_snapshot():
    specs = draw_replicas(6, seed=3, sizes=(16, 32))
    e = Ensemble(specs, precision=32)
    before = [e.conserved_sums(r) for r in range(len(specs))]
    e.step(40)
    for r, s in enumerate(specs):
        n = s.nx * s.ny * s.nz
        rho_sum, phi_sum = e.conserved_sums(r)
        assert abs(rho_sum - before[r][0]) < 1e-6 * n, r
        assert abs(phi_sum - before[r][1]) < 1e-4 * max(1.0, abs(before[r][1])) + 1e-5 * n ** 0.5, r
        dst = L.DeviceArray(4 * n, np.float32)
        e.snapshot_device(r, dst.ptr)
        got = dst.download().astype(np.float64)
        dst.free()
        phi, _, u = e.snapshot(r)
        assert_exact(got[:n], phi, f"replica {r} device phi")
        assert_exact(got[n:], u, f"replica {r} device u")
    e.close()


def test_c4_sizes_match_single_lattice():
    # the ensemble's own sizes {48, 64, 80, 96}: two replicas checked in full
    specs = draw_replicas(8, seed=11)
    by_size = {}
    for r, s in enumerate(specs):
        by_size.setdefault(s.nx, r)
    e = Ensemble(specs, precision=32)
    e.tune(reps=1)
    print(f"tuned {e.config()}", flush=True)
    e.step(20)
    for n, r in sorted(by_size.items())[:2]:
        phi, rho, u, f = single_run(specs[r], 20, 32)
        ephi, erho, eu = e.snapshot(r)
        assert_exact(ephi, phi, f"replica {r} (n={n}) phi")
        assert_exact(eu, u, f"replica {r} (n={n}) u")
    e.close()


def test_c4_full_256_replicas():
    # C4 at its stated size, the bench's workload: 256 replicas of sizes
    # {48, 64, 80, 96} (110.5 M sites) in one persistent launch per step,
    # autotuned, through the graph.  One replica of each size and the last
    # one are checked bit for bit against single-lattice runs; every
    # replica keeps its mass and its phi sum.
    specs = draw_replicas(256, seed=2026)
    e = Ensemble(specs, precision=32)
    print(f"256 replicas, {e.nsite} sites", flush=True)
    assert e.nsite == 110460928
    sums0 = [e.conserved_sums(r) for r in range(len(specs))]
    e.tune(reps=1)
    print(f"tuned {e.config()}", flush=True)
    e.step(10)
    check = {}
    for r, s in enumerate(specs):
        check.setdefault(s.nx, r)
    for r in sorted(set(check.values()) | {len(specs) - 1}):
        phi, rho, u, f = single_run(specs[r], 10, 32)
        ephi, erho, eu = e.snapshot(r)
        assert_exact(ephi, phi, f"replica {r} (n={specs[r].nx}) phi")
        assert_exact(erho, rho, f"replica {r} (n={specs[r].nx}) rho")
        assert_exact(eu, u, f"replica {r} (n={specs[r].nx}) u")
    for r, s in enumerate(specs):
        n3 = s.nx * s.ny * s.nz
        m, p = e.conserved_sums(r)
        assert abs(m - n3) <= 1e-6 * n3, f"replica {r}: mass {m} against {n3}"
        assert abs(p - sums0[r][1]) <= 1e-6 * n3, f"replica {r}: phi sum {p} against {sums0[r][1]}"
    e.close()


def test_bad_arguments_fail_loudly():
    with pytest.raises(RuntimeError, match="tile"):
        e = Ensemble(small_specs(), precision=32)
        try:
            e.set_config(48, 4, xplanes=4)
        finally:
            e.close()
    with pytest.raises(RuntimeError, match="precision"):
        Ensemble(small_specs(), precision=16)
