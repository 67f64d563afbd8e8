"""The surrogate at its stated configuration, with the trained network
shipped in the package (ludwig_amd/surrogate_unet.npz: BF16 weights, FP32
biases, the normaliser and the jump gain calibrated on training replicas).

Replicas: the 15 held-out 64^3 replicas of the 256-replica training
ensemble (20 % of seeds, never trained on) and 48 fresh 64^3 replicas drawn
from the same parameter distribution with seeds no training run has seen.
From the lattice Boltzmann state at t = 500, plain LB runs to
T = 8 (250 + 20) = 2160 steps and surrogate + LB runs 8 x (network jump of
250 steps + 20 LB correction steps), every replica advanced together, each
timed on the GPU.  L(T) from the structure factor.  The run is
deterministic (fixed-order reductions, tile shapes that change only which
block computes a site), so the errors below are reproducible to the last
digit; only the speedup varies with the machine.

Measured on MI355X: speedup 6.0-6.1x, ensemble-mean L(T) error 0.39 %, mean
per-replica |L(T) error| 2.9 % (max 10.7 %).  The per-replica 2 % target
is not met; the bounds assert what is."""

import pytest

from ludwig_amd import surrogate_study as S

pytestmark = pytest.mark.gpu


def test_c5_surrogate_gain_on_held_out_and_fresh_64():
    print("C5: 15 held-out + 48 fresh 64^3 replicas, 8 jumps x 20 correction steps", flush=True)
    res = S.gain(size=64, corrs=(20,), fresh=48)
    c = res["corr20"]
    print({k: round(v, 5) if isinstance(v, float) else v for k, v in c.items()}, flush=True)
    assert res["replicas"] == 63 and res["sizes"] == [64]
    assert c["T"] == 2160
    assert c["speedup"] >= 5.0, c
    assert c["L_err_ensemble_mean"] <= 0.02, c
    assert c["L_err_per_replica_mean"] <= 0.035, c
    assert abs(c["L_err_per_replica_signed_mean"]) <= 0.02, c
