"""The surrogate pipeline end to end, scaled down from the full study: data
from the replica ensemble, training on the GPU with the library's kernels,
a short pushforward stage (inputs the jump pipeline itself produces), and
jumps with lattice Boltzmann correction on held-out replicas.  The
trained network must beat the zero-change prediction on held-out pairs by
2x, and its jumps must leave a smaller L(t) error than identity jumps (the
method without the network)."""

import numpy as np
import pytest

from ludwig_amd.surrogate_study import run

pytestmark = pytest.mark.gpu


def test_small_surrogate_study():
    print("surrogate study: 24 replicas of 32^3, 1500 steps, narrow parameters", flush=True)
    # a narrow parameter range keeps the small data set homogeneous
    res = run(nrep=24, steps=1500, every=250, sizes=(32,), eval_size=32, iters=800, batch=4,
              lr=5e-4, jumps=2, corr=20, seed=5, t_start=500,
              replica_kw=dict(phi0=(-0.05, 0.05), kappa=(0.035, 0.045), mobility=(0.15, 0.2)),
              pf_seconds=6.0, refresh_every=100, log=lambda m: print(m, flush=True))
    assert np.isfinite(res["train_loss_last"])
    assert res["pushforward_iters"] > 0 and np.isfinite(res["pushforward_loss_last"])
    assert res["heldout_loss"] < 0.5 * res["heldout_zero_change_loss"], res
    ev = res["evaluation_corr20"]
    assert ev["T"] == 2 * 270 and ev["t_start"] == 500
    assert ev["L_err_surrogate_mean"] < ev["L_err_identity_mean"], ev
