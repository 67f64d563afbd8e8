"""FP32-shifted storage against FP64: per-field error over time.

The same binary-fluid lattice (one replica: size, parameters, seed) is
advanced in FP64 and in FP32 storage (distributions as f_i - w_i rho0,
phi and arithmetic in FP32).  Every `every` steps the FP32 fields are
compared with the FP64 ones:

    phi, rho, u   max |error| and RMS error, and the RMS error relative to
                  the RMS of the FP64 field's deviation from its mean;
    sum rho, sum phi of the FP32 run (compensated FP64 device sums) relative
                  to their initial values (conservation drift).

Also times both precisions on the same lattice.  Run as a script; prints one
JSON object per checkpoint and a summary.

Round-off control (controls=True): two more FP64 runs of the same lattice
whose initial phi differs from the reference run's by round-off only --
phi0 rounded to FP32 (the perturbation the FP32 run's own storage makes at
step 0), and phi0 moved by one FP64 ulp at every site.  Their error against
the FP64 run at each checkpoint is how far the dynamics alone amplify a
perturbation of that size, with no FP32 arithmetic at all: the floor below
which no storage of that precision can keep the phi error.
"""

import argparse
import json
import time

import numpy as np

from . import lib
from .ensemble import Ensemble, replica


def _err(a32, a64):
    d = a32 - a64
    dev = a64 - a64.mean()
    rms = float(np.sqrt(np.mean(d * d)))
    return {"max": float(np.max(np.abs(d))), "rms": rms,
            "rel_rms": rms / max(float(np.sqrt(np.mean(dev * dev))), 1e-300)}


def _controls(n, seed, kappa, mobility, phi0):
    """FP64 lattices started from the spinodal phi of (seed, phi0) perturbed
    by round-off: {"phi0_fp32_rounded": ..., "phi0_plus_1ulp": ...}."""
    from . import BinaryFluid
    out = {}
    for name in ("phi0_fp32_rounded", "phi0_plus_1ulp"):
        b = BinaryFluid((n, n, n), kappa=kappa, mobility=mobility, tau=1.0, precision=64, fatal=True)
        b.init_spinodal(seed, phi0, 0.1)
        p = b.phi_field()
        if name == "phi0_fp32_rounded":
            q = p.astype(np.float32).astype(np.float64)
        else:
            q = np.nextafter(p, np.inf)
        b.set_phi(q)
        out[name] = (b, float(np.sqrt(np.mean((q - p) ** 2))))
    return out


def run(n=64, steps=5000, every=250, seed=8361235, kappa=0.04, mobility=0.25, phi0=0.0, log=print,
        controls=False):
    spec = replica(n, seed, phi0=phi0, kappa=kappa, mobility=mobility)
    e64 = Ensemble([spec], precision=64, fatal=True)
    e32 = Ensemble([spec], precision=32, fatal=True)
    ctl = _controls(n, seed, kappa, mobility, phi0) if controls else {}
    s0 = e32.conserved_sums(0)
    rows = []
    t = 0
    while t < steps:
        k = min(every, steps - t)
        e64.step(k)
        e32.step(k)
        for b, _ in ctl.values():
            b.step(k)
        t += k
        p64, r64, u64 = e64.snapshot(0)
        p32, r32, u32 = e32.snapshot(0)
        rho_s, phi_s = e32.conserved_sums(0)
        row = {"t": t, "phi": _err(p32, p64), "rho": _err(r32, r64), "u": _err(u32, u64),
               "drift_rho": (rho_s - s0[0]) / s0[0],
               "drift_phi": (phi_s - s0[1]) / max(abs(s0[1]), 1e-300),
               "drift_phi_per_site": (phi_s - s0[1]) / n ** 3}
        for name, (b, _) in ctl.items():
            row["control_" + name] = {"phi": _err(b.phi_field(), p64)}
        rows.append(row)
        log(json.dumps(row))
    e64.close()
    e32.close()
    for b, _ in ctl.values():
        b.close()
    # timing of the two precisions on the same lattice
    times = {}
    for prec in (64, 32):
        e = Ensemble([spec], precision=prec, fatal=True)
        e.tune(reps=2)
        e.step(10)
        lib().ldw_device_synchronize()
        t0 = time.perf_counter()
        e.step(200)
        lib().ldw_device_synchronize()
        times[prec] = (time.perf_counter() - t0) / 200
        e.close()
    first_above = next((r["t"] for r in rows if r["phi"]["rel_rms"] > 1e-6), None)
    out = {}
    for name, (_, rms0) in ctl.items():
        out["control_" + name] = {
            "initial_phi_rms_perturbation": rms0,
            "phi_rel_rms_err_final": rows[-1]["control_" + name]["phi"]["rel_rms"],
            "first_t_phi_rel_rms_above_1e-6": next(
                (r["t"] for r in rows if r["control_" + name]["phi"]["rel_rms"] > 1e-6), None)}
    return {**out, "n": n, "steps": steps, "rows": rows, "ms_per_step_fp64": 1e3 * times[64],
            "ms_per_step_fp32": 1e3 * times[32], "fp32_speedup": times[64] / times[32],
            "phi_rel_rms_err_final": rows[-1]["phi"]["rel_rms"],
            "first_t_phi_rel_rms_above_1e-6": first_above}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--every", type=int, default=250)
    ap.add_argument("--controls", action="store_true",
                    help="also the FP64 runs from round-off-perturbed phi0")
    a = ap.parse_args()
    res = run(a.n, a.steps, a.every, log=lambda m: print(m, flush=True), controls=a.controls)
    print(json.dumps({k: v for k, v in res.items() if k != "rows"}))


if __name__ == "__main__":
    main()
