"""FP32-shifted storage against FP64 on one lattice (ludwig_amd/
precision_study.py): the per-field errors stay small over a few hundred
steps and the compensated sums show the FP32 run conserves mass."""

import pytest

from ludwig_amd.precision_study import run

pytestmark = pytest.mark.gpu


def test_fp32_against_fp64_short_horizon():
    print("precision study: 32^3, 400 steps", flush=True)
    res = run(n=32, steps=400, every=100, log=lambda m: None)
    rows = res["rows"]
    assert [r["t"] for r in rows] == [100, 200, 300, 400]
    for r in rows:
        assert r["phi"]["rel_rms"] < 1e-5, r
        assert r["u"]["rel_rms"] < 1e-3, r
        assert abs(r["drift_rho"]) < 1e-6, r
        assert abs(r["drift_phi_per_site"]) < 1e-6, r
    assert rows[0]["phi"]["rel_rms"] < 1e-6
