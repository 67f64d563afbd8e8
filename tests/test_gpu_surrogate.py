"""The surrogate pipeline end to end on a small ensemble: data from the
replica ensemble, training on the GPU with the library's kernels, and jumps
with lattice Boltzmann correction on held-out replicas.  The trained network
must beat the zero-change prediction on held-out pairs (loss at most 0.8 of
it) and make a smaller pointwise phi error than identity jumps (the method
without the network)."""

import numpy as np
import pytest

from ludwig_amd.surrogate_study import run

pytestmark = pytest.mark.gpu


def test_small_surrogate_study():
    print("surrogate study: 10 replicas of 16^3 / 32^3, 400 steps", flush=True)
    res = run(nrep=10, steps=500, every=50, sizes=(16, 32), eval_size=32, iters=400, batch=2,
              lr=5e-4, jumps=3, corr=10, seed=5, t_start=100, log=lambda m: print(m, flush=True))
    assert np.isfinite(res["train_loss_last"])
    assert res["heldout_loss"] < 0.8 * res["heldout_zero_change_loss"], res
    ev = res["evaluation"]
    assert ev["T"] == 3 * 60 and ev["t_start"] == 100
    assert ev["phi_rms_err_surrogate_mean"] < ev["phi_rms_err_identity_mean"], ev
