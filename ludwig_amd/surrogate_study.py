"""Surrogate + lattice Boltzmann against plain lattice Boltzmann.

Protocol:
1. An uneven-size replica ensemble (random phi0, kappa, M per replica) is
   run with snapshots of (phi, u) every `every` steps: the training data.
2. Replicas are split by seed, 80 % for training and 20 % held out.
3. The U-Net is trained on the GPU to map the snapshot at step t to the
   one at t + every, for t >= t_start.
4. On the held-out replicas of the evaluation size, from the lattice
   Boltzmann state at t_start (domains have formed; both runs get there by
   plain LB), three runs advance by the same time T = jumps * (every + corr):
     plain LB                  T steps;
     surrogate + LB            `jumps` times: one network jump (+ every
                               steps), then `corr` LB steps;
     identity + LB (baseline)  the same without the network, i.e. jumps
                               that leave the state unchanged.
   Each is judged by the domain length L(T) from the structure factor
   (ludwig_amd.analysis), per replica and for the ensemble mean, and the
   pointwise phi error against plain LB, and timed on the GPU (all held-out
   replicas advance together).  A control run of plain LB in FP64 storage
   shows how far two faithful runs drift apart on this chaotic problem.
   Surrogate variants: the predicted (phi, u) as is, and with the sums of
   phi and of each velocity component restored to their values before the
   jump (both are conserved by the dynamics).

Run as a script for the full protocol; prints one JSON object.
"""

import argparse
import json
import os
import time

import numpy as np

from . import lib
from .analysis import domain_length
from .ensemble import Ensemble, draw_replicas
from .surrogate import Dataset, Jumper, Pushforward, Trainer, UNet


def split_by_seed(specs, frac=0.2, seed=0):
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(specs))
    nval = max(1, int(round(frac * len(specs))))
    return sorted(order[nval:].tolist()), sorted(order[:nval].tolist())


def _sync():
    lib().ldw_device_synchronize()


def evaluate(net, norm, specs, every, jumps, corr, t_start=0, log=print, variants=None,
             control=True, identity=True):
    """Plain LB, surrogate + LB and identity + LB on `specs` from the LB
    state at t_start.  control=False skips the FP64 control run and
    identity=False the identity jumps."""
    T = jumps * (every + corr)
    out = {"T": T, "t_start": t_start, "replicas": len(specs)}
    plain = Ensemble(specs, precision=32, fatal=True)
    plain.tune(reps=2)
    plain.step(t_start)              # also captures the step graphs
    _sync()
    t0 = time.perf_counter()
    plain.step(T)
    _sync()
    out["plain_s"] = time.perf_counter() - t0
    ref = [plain.snapshot(r)[0] for r in range(len(specs))]
    plain.close()
    # control: the same plain run in FP64 storage -- how far two faithful
    # lattice Boltzmann runs drift apart on this chaotic problem
    if control:
        ctl = Ensemble(specs, precision=64, fatal=True)
        ctl.step(t_start + T)
        ref64 = [ctl.snapshot(r)[0] for r in range(len(specs))]
        ctl.close()

    def jumped(with_net, u_mode="pred", conserve=False):
        e = Ensemble(specs, precision=32, fatal=True)
        e.tune(reps=2)
        e.step(t_start)
        jumper = Jumper(net, norm, u_mode, conserve=conserve)
        if with_net:                 # warm the buffers and kernels once
            e2 = Ensemble(specs, precision=32, fatal=True)
            e2.step(1)
            jumper.jump(e2)
            e2.close()
        _sync()
        t0 = time.perf_counter()
        for j in range(jumps):
            if with_net:
                jumper.jump(e)
            e.step(corr)
        _sync()
        dt = time.perf_counter() - t0
        phis = [e.snapshot(r)[0] for r in range(len(specs))]
        e.close()
        return phis, dt

    if variants is None:
        variants = {"surrogate": ("pred", False), "surrogate_conserved": ("pred", True)}
    runs = {}
    for name, (u_mode, conserve) in variants.items():
        runs[name], out[name + "_s"] = jumped(True, u_mode, conserve)
    if identity:
        runs["identity"], out["identity_s"] = jumped(False)
    if control:
        runs["fp64_control"] = ref64
    names = tuple(runs)
    rows = []
    for r, s in enumerate(specs):
        shape = (s.nx, s.ny, s.nz)
        Lr = domain_length(ref[r], shape)
        row = {"replica": r, "shape": shape[0], "L_plain": Lr}
        for name, phis in runs.items():
            Ls = domain_length(phis[r], shape)
            dev = ref[r] - ref[r].mean()
            row[f"L_{name}"] = Ls
            row[f"L_err_{name}"] = abs(Ls - Lr) / Lr
            row[f"phi_rms_err_{name}"] = float(np.sqrt(np.mean((phis[r] - ref[r]) ** 2))
                                               / max(np.sqrt(np.mean(dev ** 2)), 1e-30))
        rows.append(row)
    out["per_replica"] = rows
    for k in [f"L_err_{n}" for n in names] + [f"phi_rms_err_{n}" for n in names]:
        out[k + "_mean"] = float(np.mean([row[k] for row in rows]))
        out[k + "_max"] = float(np.max([row[k] for row in rows]))
    # error of the ensemble-mean domain length
    Lm = float(np.mean([row["L_plain"] for row in rows]))
    for name in names:
        out[f"meanL_err_{name}"] = abs(float(np.mean([row[f"L_{name}"] for row in rows])) - Lm) / Lm
    for name in variants:
        out["speedup" + name[len("surrogate"):]] = out["plain_s"] / out[name + "_s"]
    log(f"evaluation: {json.dumps({k: v for k, v in out.items() if k != 'per_replica'})}")
    return out


# The trained network shipped with the package: BF16 weights, FP32 biases
# and the normaliser, from run(..., save_to=...) (save_inference).
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "surrogate_unet.npz")


def heldout_specs(nrep=256, seed=2026, size=None):
    """The held-out 20 % of the training ensemble's replicas (split by seed,
    as run() splits them), optionally only those of one lattice size."""
    specs = draw_replicas(nrep, seed=seed)
    _, val_idx = split_by_seed(specs, 0.2, seed)
    return [specs[i] for i in val_idx if size is None or specs[i].nx == size]


FRESH_SEED = 424242


def fresh_specs(nrep, size=64, train_nrep=256, train_seed=2026):
    """nrep replicas of one lattice size drawn from the training ensemble's
    parameter distribution with a generator of their own: seeds and
    parameters that no training run has seen (checked against the
    training ensemble's seeds)."""
    specs = draw_replicas(nrep, seed=FRESH_SEED, sizes=(size,))
    seen = {s.seed for s in draw_replicas(train_nrep, seed=train_seed)}
    if any(s.seed in seen for s in specs):
        raise ValueError("a fresh replica's seed is one of the training ensemble's")
    return specs


def train_specs(nrep=256, seed=2026, size=64):
    """The training 80 % of the training ensemble's replicas, optionally
    only those of one lattice size."""
    specs = draw_replicas(nrep, seed=seed)
    tr_idx, _ = split_by_seed(specs, 0.2, seed)
    return [specs[i] for i in tr_idx if size is None or specs[i].nx == size]


def calibrate(weights, gains=(1.0, 1.05, 1.1, 1.15, 1.2), size=64, corr=20, every=250, jumps=8,
              t_start=500, max_replicas=32, conserve=True, u_too=False, log=print):
    """The jump gain (Normaliser.jump_gain_phi; with u_too also the
    velocity's) that removes the mean L(T) bias of the jump pipeline on
    TRAINING replicas of one size: each gain's signed mean L error, mean
    |L error| and ensemble-mean L error; returns (best gain, rows)."""
    net = UNet(seed=0)
    norm = net.load(weights)
    specs = train_specs(size=size)[:max_replicas]
    rows = []
    for g in gains:
        nm = norm.with_gains(g, g if u_too else 1.0)
        ev = evaluate(net, nm, specs, every, jumps, corr, t_start=t_start, log=lambda m: None,
                      variants={"surrogate": ("pred", conserve)}, control=False, identity=False)
        signed = float(np.mean([(r["L_surrogate"] - r["L_plain"]) / r["L_plain"] for r in ev["per_replica"]]))
        rows.append({"gain": g, "signed_mean": signed, "abs_mean": ev["L_err_surrogate_mean"],
                     "ensemble_mean": ev["meanL_err_surrogate"], "speedup": ev["speedup"]})
        log(f"calibration gain {g}: {rows[-1]}")
    best = min(rows, key=lambda r: abs(r["signed_mean"]))["gain"]
    return best, rows


def write_gains(src, dst, phi=1.0, u=1.0):
    """A copy of the weights file src with the jump gains set."""
    d = dict(np.load(src))
    d["norm_jump_gain_phi"] = np.array(float(phi))
    d["norm_jump_gain_u"] = np.array(float(u))
    np.savez_compressed(dst, **d)


def gain(weights=WEIGHTS, size=64, corrs=(20, 30), every=250, jumps=8, t_start=500,
         conserve=True, log=None, fresh=0):
    """Surrogate + LB against plain LB on the held-out replicas (all sizes
    with size=None), plus `fresh` never-seen replicas of that size: for each
    number of correction steps, the time-to-T speedup and the L(T) error
    per replica (mean, max, signed mean) and of the ensemble mean, with
    T = jumps * (every + corr)."""
    net = UNet(seed=0)
    norm = net.load(weights)
    if norm is None:
        raise ValueError(f"{weights}: no normaliser stored with the weights")
    specs = heldout_specs(size=size)
    if fresh:
        specs = specs + fresh_specs(fresh, size=size or 64)
    out = {"weights": os.path.basename(weights), "replicas": len(specs), "fresh": fresh,
           "sizes": sorted({s.nx for s in specs}), "jumps": jumps, "every": every, "t_start": t_start,
           "sums_restored": conserve}
    for c in corrs:
        ev = evaluate(net, norm, specs, every, jumps, c, t_start=t_start,
                      log=log or (lambda m: None), variants={"surrogate": ("pred", conserve)},
                      control=False, identity=False)
        signed = [(r["L_surrogate"] - r["L_plain"]) / r["L_plain"] for r in ev["per_replica"]]
        out[f"corr{c}"] = {"T": ev["T"], "speedup": ev["speedup"], "plain_s": ev["plain_s"],
                           "surrogate_s": ev["surrogate_s"],
                           "L_err_per_replica_mean": ev["L_err_surrogate_mean"],
                           "L_err_per_replica_max": ev["L_err_surrogate_max"],
                           "L_err_per_replica_signed_mean": float(np.mean(signed)),
                           "L_err_ensemble_mean": ev["meanL_err_surrogate"],
                           "phi_rms_err_mean": ev["phi_rms_err_surrogate_mean"]}
    return out


def run(nrep=256, steps=4000, every=250, sizes=(48, 64, 80, 96), eval_size=0, iters=2000,
        batch=4, lr=5e-4, jumps=8, corr=30, seed=2026, max_eval=None, train_seconds=None,
        t_start=500, replica_kw=None, init_from=None, save_to=None, log=print, pf_seconds=0.0,
        pf_lr=None, p_pf=0.5, refresh_every=400, eval_corrs=None, augment=False):
    """Data, training and the held-out evaluation.  pf_seconds > 0 adds a
    second training stage on pushforward pairs (Trainer.fit_pushforward):
    the inputs the jump pipeline with `corr` correction steps reaches after
    one network jump, against the true state one jump later."""
    specs = draw_replicas(nrep, seed=seed, sizes=sizes, **(replica_kw or {}))
    train_idx, val_idx = split_by_seed(specs, 0.2, seed)
    t0 = time.perf_counter()
    ens = Ensemble(specs, precision=32, fatal=True)
    ens.tune(reps=2)
    data = Dataset(ens, steps // every, every, offset=corr if pf_seconds else 0)
    ens.close()
    _sync()
    t_data = time.perf_counter() - t0
    log(f"data: {nrep} replicas, {steps} steps, snapshots every {every}: {t_data:.1f} s")
    kmin = (t_start + every - 1) // every
    norm = data.normaliser(sample=min(32, len(data.pairs(kmin=kmin))), kmin=kmin)
    net = UNet(seed=seed)
    if init_from:
        stored = net.load(init_from)
        if stored is not None:
            norm = stored
        log(f"weights from {init_from} (after {net.step_count} steps)")
    tr = Trainer(net, norm, batch=batch, augment=augment, seed=seed)
    tpairs = [(r, k) for (r, k) in data.pairs(kmin=kmin) if r in set(train_idx)]
    vpairs = [(r, k) for (r, k) in data.pairs(kmin=kmin) if r in set(val_idx)]
    t0 = time.perf_counter()
    losses = []
    if iters > 0 and (train_seconds is None or train_seconds > 0):
        losses = tr.fit(data, tpairs, iters, lr=lr, lr_final=lr * 0.1, log=log, max_seconds=train_seconds)
    _sync()
    t_train = time.perf_counter() - t0
    losses_pf = []
    t_pf = 0.0
    if pf_seconds:
        if save_to:
            net.save(save_to[:-4] + "_stage1.npz" if save_to.endswith(".npz") else save_to + "_stage1", norm)
        t0 = time.perf_counter()
        pf = Pushforward(data, train_idx, kmin, conserve=True)
        plr = pf_lr if pf_lr is not None else lr * 0.3
        l1, losses_pf = tr.fit_pushforward(data, pf, tpairs, 10 ** 9, lr=plr, lr_final=plr * 0.1, p_pf=p_pf,
                                           refresh_every=refresh_every, log=log, max_seconds=pf_seconds)
        losses += l1
        pf.free()
        _sync()
        t_pf = time.perf_counter() - t0
    vsub = vpairs if len(vpairs) <= 64 else [vpairs[i] for i in
                                              np.random.default_rng(1).choice(len(vpairs), 64, replace=False)]
    val_loss = tr.evaluate(data, vsub)
    zero_loss = tr.evaluate(data, vsub, zero=True)
    last = float(np.mean(losses[-20:])) if losses else float("nan")
    log(f"training: {len(losses)} plain + {len(losses_pf)} pushforward iterations in "
        f"{t_train + t_pf:.1f} s, last loss {last:.4e}, "
        f"held-out loss {val_loss:.4e} (zero-change prediction {zero_loss:.4e})")
    data.free()
    if save_to:
        net.save(save_to, norm)
        net.save_inference(save_to[:-4] + "_inference.npz" if save_to.endswith(".npz")
                           else save_to + "_inference", norm)
    # eval_size 0: every held-out replica
    ev_specs = [specs[i] for i in val_idx if specs[i].nx == eval_size or not eval_size] \
        or [specs[i] for i in val_idx]
    if max_eval:
        ev_specs = ev_specs[:max_eval]
    result = {"replicas": nrep, "steps": steps, "every": every, "data_s": t_data,
              "train_iters": len(losses), "train_s": t_train, "train_loss_last": last,
              "pushforward_iters": len(losses_pf), "pushforward_s": t_pf,
              "pushforward_loss_last": float(np.mean(losses_pf[-20:])) if losses_pf else None,
              "heldout_loss": val_loss, "heldout_zero_change_loss": zero_loss,
              "normaliser": norm.to_dict()}
    if ev_specs:
        for c in (eval_corrs or (corr,)):
            result[f"evaluation_corr{c}"] = evaluate(net, norm, ev_specs, every, jumps, c,
                                                     t_start=kmin * every, log=log)
    return result


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--every", type=int, default=250)
    ap.add_argument("--eval-size", type=int, default=0, help="0: every held-out size")
    ap.add_argument("--iters", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--jumps", type=int, default=8)
    ap.add_argument("--corr", type=int, default=30)
    ap.add_argument("--max-eval", type=int, default=0)
    ap.add_argument("--train-seconds", type=float, default=0.0)
    ap.add_argument("--t-start", type=int, default=500)
    ap.add_argument("--init-from", default=None, help="weights (.npz) to start from")
    ap.add_argument("--save-to", default=None, help="write the trained weights (.npz) here")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--pf-seconds", type=float, default=0.0,
                    help="length of the pushforward training stage (0: none)")
    ap.add_argument("--pf-lr", type=float, default=None)
    ap.add_argument("--p-pf", type=float, default=0.5, help="share of pushforward batches in that stage")
    ap.add_argument("--refresh-every", type=int, default=400)
    ap.add_argument("--eval-corr", type=int, nargs="*", default=None,
                    help="correction steps of the evaluations (default: --corr)")
    ap.add_argument("--augment", action="store_true",
                    help="map training pairs by random elements of the cubic group")
    a = ap.parse_args()
    res = run(a.replicas, a.steps, a.every, eval_size=a.eval_size, iters=a.iters, batch=a.batch, lr=a.lr,
              jumps=a.jumps, corr=a.corr, seed=a.seed, max_eval=a.max_eval or None,
              train_seconds=a.train_seconds, t_start=a.t_start, init_from=a.init_from,
              save_to=a.save_to, pf_seconds=a.pf_seconds, pf_lr=a.pf_lr, p_pf=a.p_pf,
              refresh_every=a.refresh_every, eval_corrs=a.eval_corr, augment=a.augment,
              log=lambda m: print(m, flush=True))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
