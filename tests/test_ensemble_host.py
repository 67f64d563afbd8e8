"""CPU checks of the replica-ensemble host side: the ctypes mirror of
ldw_replica_t matches the C header's layout, and the random ensemble
generator keeps its ranges and the explicit scheme's stability bound."""

import ctypes
import os
import subprocess

import numpy as np
import pytest

from ludwig_amd.ensemble import ReplicaSpec, draw_replicas, max_stable_mobility, replica

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_replica_struct_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "ludwig_hip.h"\n'
        "int main(void) {\n"
        '  printf("%zu", sizeof(ldw_replica_t));\n'
        + "".join(f'  printf(" %zu", offsetof(ldw_replica_t, {f}));\n'
                  for f, _ in ReplicaSpec._fields_)
        + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    want = [ctypes.sizeof(ReplicaSpec)] + [getattr(ReplicaSpec, f).offset for f, _ in ReplicaSpec._fields_]
    assert got == want


def test_stability_bound_is_the_linear_amplification_limit():
    # 1 - M k2 (a + 3 b phi^2 + kappa k2) = -1 at the bound, k2 = 12
    a, b, kappa, phimax = -0.0625, 0.0625, 0.04, 1.1
    m = max_stable_mobility(a, b, kappa, phimax)
    assert 1 - m * 12 * (a + 3 * b * phimax**2 + kappa * 12) == pytest.approx(-1.0, abs=1e-14)
    # the stated M = 0.45 at kappa = 0.04 is beyond it, 0.25 is inside
    assert 0.25 < m < 0.45


def test_draw_replicas_ranges():
    reps = draw_replicas(256, seed=7)
    assert len(reps) == 256
    sizes = {r.nx for r in reps}
    assert sizes <= {48, 64, 80, 96} and len(sizes) == 4
    for r in reps:
        assert r.nx == r.ny == r.nz
        assert -0.3 <= r.phi0 <= 0.3
        assert 0.02 <= r.kappa <= 0.08
        assert 0.1 <= r.mobility <= 0.5
        assert r.mobility <= 0.95 * max_stable_mobility(r.a, r.b, r.kappa) + 1e-15
        assert r.eta_shear == pytest.approx(1.0 / 6.0)
    seeds = [r.seed for r in reps]
    assert len(set(seeds)) == len(seeds)
    # deterministic
    again = draw_replicas(256, seed=7)
    assert [(r.nx, r.seed, r.kappa) for r in again] == [(r.nx, r.seed, r.kappa) for r in reps]
    nsite = sum(r.nx ** 3 for r in reps)
    assert 0.5 * 256 * 442_000 < nsite < 1.5 * 256 * 442_000


def test_replica_defaults():
    r = replica((16, 20, 24), 5)
    assert (r.nx, r.ny, r.nz, r.seed) == (16, 20, 24, 5)
    assert r.eta_shear == r.eta_bulk == pytest.approx(1.0 / 6.0)
