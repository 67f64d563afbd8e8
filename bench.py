#!/usr/bin/env python3
"""Benchmark of the D3Q19 binary-fluid step on one MI355X.

Workload (a step = one full time step of the path over the whole lattice):
symmetric free energy + finite-difference Cahn-Hilliard + D3Q19 MRT, the
plain binary-fluid configuration: 128^3 periodic, FP64, A = -0.0625,
B = 0.0625, kappa = 0.04, M = 0.25 (the stated 0.45 is unstable for the explicit
Cahn-Hilliard step, in the reference too), tau = 1, phi0 = 0 with +-0.05 spinodal
noise, rho = 1, u = 0.  Inputs are resident in HBM before the timed region.

Prints one JSON line:
  value        MLUPS = sites * steps / wall time of the timed region
  roofline     for the dominant kernel (binary_step_kernel, one launch per
               step): algorithmic bytes per launch (376 B/site in FP64) /
               its average duration, from HIP events on the library's stream
               around the timed region divided by the launches in it; peak
               8000 GB/s; traffic = HBM bytes per launch measured with
               rocprofv3 FETCH_SIZE/WRITE_SIZE counters
  cpu_baseline the reference's own serial/OpenMP build (oracle/_ref/Ludwig.exe,
               compiled from the reference sources) on the host cores: 100
               steps at 128^3 and 20 at 256^3; the CPU oracle when that binary
               is absent
  extra        each over at least 100 steps: 256^3 FP64 (with its own
               roofline block) and FP32-shifted, 128^3 FP32; the binary fluid
               at 128^3 with two Lees-Edwards planes; the 256-replica
               uneven-size ensemble (FP32); the surrogate: time-to-T speedup
               and L(T) error of surrogate + LB against plain LB on the
               held-out replicas, and the U-Net's BF16 MFMA rate

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N]
"""

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
BYTES_SITE = {64: {"collide": 360, "phi_force_mu": 40, "cahn_hilliard": 48, "force_ch": 64, "binary_step": 376, "step": 448},
              32: {"collide": 180, "phi_force_mu": 20, "cahn_hilliard": 24, "force_ch": 32, "binary_step": 188, "step": 224}}
# C2 with M = 0.25: at the stated M = 0.45 the explicit Cahn-Hilliard step is
# unstable (the reference reaches NaN by step ~30; tests/golden/binary32_m045_*)
KERNEL_DESC = {
    "binary_step": "binary_step_kernel (force + Cahn-Hilliard + MRT collision + pull "
                   "propagation in one pass)",
    "collide": "collide_kernel (MRT collision + pull propagation)",
    "force_ch": "force_ch_kernel (force + Cahn-Hilliard)",
}
# HBM bytes per site of binary_step_kernel (FP64, default 64 x 8 tiles) from
# rocprofv3 FETCH_SIZE and WRITE_SIZE, each scaled by the same counter's
# reading for the plain copy kernel (fetch x 1.9885, write x 0.9701);
# profiles/binary_step_hbm_traffic_pmc.md
TRAFFIC_B_SITE = {("binary_step", 64, 128): 388.4, ("binary_step", 64, 256): 421.0}
C2 = dict(a=-0.0625, b=0.0625, kappa=0.04, mobility=0.25, tau=1.0, order=1)
SEED = 8361235


def _elapsed(lib, a, b):
    """Milliseconds between stream marks a and b (HIP events)."""
    import ctypes
    import ludwig_amd as L
    ms = ctypes.c_double()
    L._ck(lib.ldw_stream_elapsed(a, b, ctypes.byref(ms)), "ldw_stream_elapsed")
    return ms.value


def run_gpu(n, steps, warmup, precision, breakdown_steps=20):
    """`steps` timed steps after `warmup`: wall time, and HIP events on the
    library's stream around the same region; then a per-kernel breakdown
    (events around every launch) over `breakdown_steps` more steps."""
    import ludwig_amd as L
    b = L.BinaryFluid((n, n, n), precision=precision, fatal=True, **C2)
    b.init_spinodal(SEED, 0.0, 0.1)
    lib = L.lib()
    b.step(warmup)
    b.synchronize()
    t0 = time.perf_counter()
    lib.ldw_stream_mark(0)
    b.step(steps)
    lib.ldw_stream_mark(1)
    b.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    event_ms = _elapsed(lib, 0, 1)
    # kernel durations: HIP events around each launch, on the library stream
    lib.ldw_timer_reset()
    lib.ldw_timer_enable(1)
    b.step(breakdown_steps)
    b.synchronize()
    lib.ldw_timer_enable(0)
    kt, kc = {}, {}
    import ctypes
    for name in ("binary_step", "collide", "force_ch", "phi_force_mu", "cahn_hilliard", "mu", "propagate"):
        ms = ctypes.c_double()
        cnt = ctypes.c_int()
        lib.ldw_timer_read(name.encode(), ctypes.byref(ms), ctypes.byref(cnt))
        if cnt.value:
            kt[name] = ms.value / cnt.value
            kc[name] = cnt.value
    r, p = b.conserved_sums()
    ok = bool(abs(r - n ** 3) / n ** 3 < 1e-6)
    b.close()
    nsite = n ** 3
    mlups = nsite * steps / wall / 1e6
    # one launch per step of a single kernel family: the region's event time
    # per step is that kernel's average launch duration
    single = len(kc) == 1 and list(kc.values())[0] == breakdown_steps
    return {"n": n, "precision": precision, "steps": steps, "wall_s": wall,
            "ms_per_step": 1e3 * wall / steps, "event_ms_per_step": event_ms / steps, "mlups": mlups,
            "hbm_frac_step": mlups * 1e6 * BYTES_SITE[precision]["step"] / (HBM_PEAK_GBS * 1e9),
            "kernel_ms": kt, "single_kernel": list(kc)[0] if single else None, "mass_ok": ok}


def roofline(r, traffic_b_site=None):
    """The roofline block of a run_gpu() result: the dominant kernel, its
    algorithmic bytes per launch over its average launch duration."""
    kms = r["kernel_ms"]
    if not kms:
        return None
    name = r["single_kernel"] or max(kms, key=kms.get)
    # a single kernel per step: its launches fill the timed region, whose
    # event time per step is the average launch duration (it cannot exceed
    # the wall time per step); otherwise the per-launch events
    ms = r["event_ms_per_step"] if r["single_kernel"] else kms[name]
    nsite = r["n"] ** 3
    bps = BYTES_SITE[r["precision"]][name]
    achieved = bps * nsite / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_b_site * nsite if traffic_b_site else None,
            "kernel": KERNEL_DESC.get(name, name),
            "algorithmic_bytes_per_site": bps, "avg_kernel_ms": ms,
            "timing": ("HIP events on the library stream around the timed region / launches"
                       if r["single_kernel"] else "HIP events around each launch"),
            "launches": r["steps"] if r["single_kernel"] else None}


def run_ensemble(nrep, steps, warmup, precision=32, seed=2026):
    """The uneven-replica ensemble: nrep cubic replicas with sizes from
    {48, 64, 80, 96}, phi0, kappa, M drawn per replica; autotuned persistent
    kernel, one graph launch per step."""
    from ludwig_amd.ensemble import Ensemble, draw_replicas
    specs = draw_replicas(nrep, seed=seed)
    t0 = time.perf_counter()
    e = Ensemble(specs, precision=precision, fatal=True)
    t_init = time.perf_counter() - t0
    rows = e.tune(reps=3)
    cfg = e.config()
    e.step(warmup)
    import ludwig_amd as L
    L.lib().ldw_device_synchronize()
    t0 = time.perf_counter()
    e.step(steps)
    L.lib().ldw_device_synchronize()
    wall = time.perf_counter() - t0
    nsite = e.nsite
    mass = [e.conserved_sums(r)[0] for r in (0, nrep - 1)]
    e.close()
    mlups = nsite * steps / wall / 1e6
    bps = BYTES_SITE[precision]["binary_step"]
    return {"replicas": nrep, "sites": nsite, "precision": precision, "steps": steps,
            "ms_per_step": 1e3 * wall / steps, "mlups": mlups,
            "hbm_GBps_algorithmic": mlups * 1e6 * bps / 1e9,
            "hbm_frac_step": mlups * 1e6 * BYTES_SITE[precision]["step"] / (HBM_PEAK_GBS * 1e9),
            "tuned": cfg, "tuner_trials": rows.tolist(), "init_s": t_init,
            "mass_first_last": mass}


def run_conserve2(n=128, steps=100, warmup=5):
    """cahn_hilliard_options_conserve 2 at n^3 FP64: ms per step with the
    default compensated tree sum and with the one-thread site-order sum
    (the reference's one-thread order, bit-identical with it)."""
    import ludwig_amd as L
    lib = L.lib()
    out = {"n": n, "steps": steps}
    for mode, key in ((0, "tree_sum"), (1, "site_order_sum")):
        lib.ldw_conserve2_site_order(mode)
        b = L.BinaryFluid((n, n, n), precision=64, fatal=True, conserve=2, **C2)
        b.init_spinodal(SEED, 0.0, 0.1)
        _, p0 = b.conserved_sums()
        b.step(warmup)
        b.synchronize()
        nst = steps if mode == 0 else max(10, steps // 10)
        t0 = time.perf_counter()
        b.step(nst)
        b.synchronize()
        out[key + "_ms_per_step"] = 1e3 * (time.perf_counter() - t0) / nst
        _, p1 = b.conserved_sums()
        out[key + "_phi_sum_drift"] = abs(p1 - p0)
        b.close()
    lib.ldw_conserve2_site_order(0)
    return out


# Lees-Edwards path (FP64).  The single-pass step (two launches of the
# tiled kernel, timer "le_step") reads phi, u and f and writes phi, rho, u,
# force and f: 400 bytes per site, the plane-free step's 376 plus the stored
# force.  The staged path (LDW_LE_FUSED=0) has one kernel per stage, with the
# algorithmic bytes per site of each below.  The plane-only kernels (buffer
# planes, faces, sums, LB plane conditions) touch O(N^2) sites and are left
# out of both.
LE_STEP_BYTES_SITE = 400
LE_BYTES_SITE = {"le_gradient": 40,        # phi in; grad (3) and delsq out
                 "le_force": 64,           # phi, grad, delsq in; force out
                 "le_cahn_hilliard": 48,   # phi, delsq, u in; phi out
                 "collide": 360}           # f in and out, force in, rho and u out


def run_le(n=128, steps=50, warmup=5, nplanes=2, uy=0.05):
    """The binary fluid with Lees-Edwards planes (steady shear from the
    shear-profile start): whole-step MLUPS and each kernel's HBM rate."""
    import ctypes
    import ludwig_amd as L
    b = L.BinaryFluid((n, n, n), fatal=True, le_planes=nplanes, le_uy=uy, **C2)
    b.init_spinodal(SEED, 0.0, 0.1)
    b.le_init_shear_profile()
    lib = L.lib()
    b.step(warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    b.synchronize()
    wall = time.perf_counter() - t0
    lib.ldw_timer_reset()
    lib.ldw_timer_enable(1)
    b.step(min(steps, 20))
    b.synchronize()
    lib.ldw_timer_enable(0)
    kt, rate = {}, {}
    bytes_site = dict(LE_BYTES_SITE, le_step=LE_STEP_BYTES_SITE)
    for name in ("le_planes", "le_step", "le_gradient", "le_force", "le_cahn_hilliard", "collide",
                 "le_lb_bc"):
        ms, cnt = ctypes.c_double(), ctypes.c_int()
        lib.ldw_timer_read(name.encode(), ctypes.byref(ms), ctypes.byref(cnt))
        if cnt.value:
            kt[name] = ms.value / cnt.value
            if name in bytes_site:
                rate[name] = bytes_site[name] * n ** 3 / (kt[name] * 1e-3) / 1e9
    r, _ = b.conserved_sums()
    b.close()
    mlups = n ** 3 * steps / wall / 1e6
    bps = LE_STEP_BYTES_SITE if "le_step" in kt else sum(LE_BYTES_SITE.values())
    return {"n": n, "planes": nplanes, "plane_speed": uy, "steps": steps,
            "ms_per_step": 1e3 * wall / steps, "mlups": mlups,
            "path": "single-pass" if "le_step" in kt else "staged",
            "bytes_per_site_algorithmic": bps,
            "hbm_frac_step": mlups * 1e6 * bps / (HBM_PEAK_GBS * 1e9),
            "kernel_ms": kt, "kernel_GBps": rate,
            "mass_ok": bool(abs(r - n ** 3) / n ** 3 < 1e-6)}


BF16_PEAK_TFLOPS = 2500.0      # MI355X dense BF16 MFMA


def run_surrogate_conv(n=64, batch=8, reps=5):
    """Inference of the surrogate U-Net on batch x n^3: MFMA FLOPs of its 12
    convolutions / wall time, and the largest layer (d0, 80 -> 32 channels at
    full resolution) timed alone."""
    import ctypes
    import numpy as np
    import ludwig_amd as L
    from ludwig_amd.surrogate import LAYERS, LEVEL, UNet
    net = UNet(seed=1)
    shape = (n, n, n)
    npos = batch * n ** 3
    x = L.DeviceArray.from_host(np.zeros(npos * 16, np.uint16), np.uint16)
    net.forward(x, batch, shape, keep=False)
    L.lib().ldw_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        net.forward(x, batch, shape, keep=False)
    L.lib().ldw_device_synchronize()
    dt = (time.perf_counter() - t0) / reps
    flops = sum(2 * 27 * ci * co * npos / 8 ** LEVEL[name] for name, ci, co in LAYERS)
    b = net._bufs[(batch,) + shape]
    lib = L.lib()
    args = (ctypes.c_void_p(b["cat0"].ptr), ctypes.c_void_p(net.wf["d0"].ptr), net.bptr("d0"), None,
            ctypes.c_void_p(b["a_d0"].ptr), batch, n, n, n, 80, 32)
    lib.ldw_conv3d_fwd_silu(*args)
    lib.ldw_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps * 4):
        lib.ldw_conv3d_fwd_silu(*args)
    lib.ldw_device_synchronize()
    dl = (time.perf_counter() - t0) / (reps * 4)
    fl = 2 * 27 * 80 * 32 * npos
    x.free()
    return {"lattice": n, "batch": batch, "forward_ms": 1e3 * dt, "forward_tflops": flops / dt / 1e12,
            "forward_frac_bf16_peak": flops / dt / 1e12 / BF16_PEAK_TFLOPS,
            "d0_layer_ms": 1e3 * dl, "d0_layer_tflops": fl / dl / 1e12,
            "d0_layer_frac_bf16_peak": fl / dl / 1e12 / BF16_PEAK_TFLOPS}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fp:
            for line in fp:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _ludwig_exe(exe, n, steps, threads):
    """One run of the reference's Ludwig.exe on the C2 input at n^3:
    (seconds in its 'Time step loop' timer, steps timed)."""
    inp = f"""N_cycles {steps}
size {n}_{n}_{n}
grid 1_1_1
viscosity 0.16666666666666667
free_energy symmetric
symmetric_a -0.0625
symmetric_b 0.0625
symmetric_kappa 0.04
mobility 0.25
phi0 0.0
phi_initialisation spinodal
noise 0.1
random_seed {SEED}
fd_gradient_calculation 3d_7pt_fluid
fd_advection_scheme_order 1
colloid_init no_colloids
periodicity 1_1_1
freq_statistics {steps}
config_at_end no
"""
    with tempfile.TemporaryDirectory() as tmp:
        with open(os.path.join(tmp, "input"), "w") as fp:
            fp.write(inp)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
                   OMP_PLACES="cores")
        out = subprocess.run([exe], cwd=tmp, env=env, capture_output=True, text=True,
                             timeout=900).stdout
    m = re.search(r"Time step loop:\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+\((\d+) calls\)", out)
    if not m:
        raise RuntimeError(f"Ludwig.exe {n}^3: no 'Time step loop' timer in its output")
    return float(m.group(3)), int(m.group(5))


def cpu_baseline(runs=((128, 100), (256, 20))):
    """Time the reference's CPU build on this host's cores: the C2 input at
    each (size, steps) of `runs`; the value is the first run's MLUPS."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    exe = os.path.join(HERE, "oracle", "_ref", "Ludwig.exe")
    cpu = _cpu_model()
    if os.path.exists(exe):
        rows = []
        for n, steps in runs:
            t, calls = _ludwig_exe(exe, n, steps, threads)
            rows.append({"n": n, "steps": calls, "loop_s": t, "mlups": n ** 3 * calls / t / 1e6})
            print(f"# cpu baseline {n}^3: {rows[-1]}", file=sys.stderr, flush=True)
        return {"value": rows[0]["mlups"], "unit": "MLUPS", "cores": threads, "kind": "reference",
                "cpu_model": cpu, "runs": rows,
                "sample": "reference Ludwig.exe (gcc -O2 -fopenmp -DNDEBUG), binary fluid "
                          + ", ".join(f"{r['n']}^3 x {r['steps']} steps" for r in rows)
                          + f", 'Time step loop' timer, {threads} OpenMP threads"
                            " (OMP_PROC_BIND=close, OMP_PLACES=cores)"}
    # fall back to the C restatement
    import oracle as O
    n, steps = runs[0][0], 10
    shape = (n, n, n)
    f0 = O.equilibrium(shape, rho_uniform=1.0)
    phi0 = O.spinodal(shape, SEED, 0.0, 0.1)
    t0 = time.perf_counter()
    O.binary_steps(shape, steps, O.lb_param(tau=1.0), O.fe_param(), f0, phi0)
    dt = time.perf_counter() - t0
    return {"value": n ** 3 * steps / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "cpu_model": cpu,
            "sample": f"C oracle (gcc -O2 -fopenmp), {n}^3 binary fluid, {steps} steps"}


def run_surrogate_gain():
    """Surrogate + LB against plain LB to the same time T with the shipped
    weights: on 64^3 (the stated configuration) the 15 held-out replicas of
    the training ensemble plus 48 fresh never-trained ones, with 20 (the
    stated count) and 30 correction steps; and every held-out size with 20."""
    from ludwig_amd import surrogate_study as S
    if not os.path.exists(S.WEIGHTS):
        return {"status": f"{os.path.basename(S.WEIGHTS)} absent"}
    out = {"held_out_and_fresh_64^3": S.gain(size=64, corrs=(20, 30), fresh=48)}
    print(f"# surrogate 64^3: {out['held_out_and_fresh_64^3']}", file=sys.stderr, flush=True)
    out["held_out_all_sizes"] = S.gain(size=None, corrs=(20,))
    print(f"# surrogate all sizes: {out['held_out_all_sizes']}", file=sys.stderr, flush=True)
    return out


def _surrogate_summary(g):
    """The metric's second half at the stated configuration (64^3, 8 jumps
    x 20 correction steps): time-to-T speedup and L(T) error of surrogate +
    LB against plain LB."""
    if not g or "held_out_and_fresh_64^3" not in g:
        return None
    r = g["held_out_and_fresh_64^3"]
    c = r["corr20"]
    return {"config": f"64^3, {r['replicas']} replicas ({r['replicas'] - r['fresh']} held out of the "
                      f"training ensemble, {r['fresh']} fresh seeds), 8 jumps x (250 + 20) steps from t = 500",
            "speedup": c["speedup"], "L_err_ensemble_mean": c["L_err_ensemble_mean"],
            "L_err_per_replica_mean": c["L_err_per_replica_mean"],
            "L_err_per_replica_signed_mean": c["L_err_per_replica_signed_mean"],
            "L_err_per_replica_max": c["L_err_per_replica_max"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--profile-run", type=int, default=0,
                    help="only run this many steps of the main workload (for rocprofv3)")
    ap.add_argument("--precision", type=int, default=64)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the plain copy kernel (1 GiB, 20 launches)")
    ap.add_argument("--profile-ensemble", type=int, default=0,
                    help="only run this many steps of the 256-replica ensemble (for rocprofv3)")
    ap.add_argument("--profile-le", type=int, default=0,
                    help="only run this many steps of the Lees-Edwards workload (for rocprofv3)")
    ap.add_argument("--profile-surrogate", type=int, default=0,
                    help="only run this many U-Net inferences on 8 x 64^3 (for rocprofv3)")
    args = ap.parse_args()
    if args.profile_ensemble:
        from ludwig_amd.ensemble import Ensemble, draw_replicas
        e = Ensemble(draw_replicas(256, seed=2026), precision=32, fatal=True)
        e.tune(reps=2)
        e.step(args.profile_ensemble)
        e.close()
        print(f"profile run: ensemble of 256 replicas, {args.profile_ensemble} steps")
        return
    if args.profile_le:
        r = run_le(steps=args.profile_le, warmup=1)
        print(f"profile run: lees-edwards {r}")
        return
    if args.profile_surrogate:
        r = run_surrogate_conv(reps=args.profile_surrogate)
        print(f"profile run: surrogate inference {r}")
        return
    if args.copy_calibration:
        import ctypes
        import ludwig_amd as L
        ms = ctypes.c_double()
        nbytes = 1 << 30
        L._ck(L.lib().ldw_copy_calibration(ctypes.c_longlong(nbytes), 20, ctypes.byref(ms)),
              "ldw_copy_calibration")
        print(json.dumps({"copy_bytes": nbytes, "ms": ms.value,
                          "GBps_read_plus_write": 2 * nbytes / (ms.value * 1e-3) / 1e9}))
        return
    if args.profile_run:
        import ludwig_amd as L
        n = args.size
        b = L.BinaryFluid((n, n, n), precision=args.precision, fatal=True, **C2)
        b.init_spinodal(SEED, 0.0, 0.1)
        b.step(args.profile_run)
        b.synchronize()
        b.close()
        print(f"profile run: {n}^3 fp{args.precision}, {args.profile_run} steps")
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    xsteps = max(100, args.steps)          # every extra over >= 100 steps
    main_run = run_gpu(args.size, args.steps, args.warmup, 64)
    print(f"# main: {main_run}", file=sys.stderr, flush=True)
    extra = {}
    if not args.no_extra:
        for n, prec in ((256, 64), (256, 32), (128, 32)):
            r = run_gpu(n, xsteps, 10, prec)
            extra[f"{n}^3_fp{prec}"] = {k: r[k] for k in ("steps", "mlups", "ms_per_step",
                                                          "event_ms_per_step", "hbm_frac_step",
                                                          "kernel_ms", "mass_ok")}
            extra[f"{n}^3_fp{prec}"]["roofline"] = roofline(
                r, TRAFFIC_B_SITE.get(("binary_step", prec, n)))
            print(f"# extra {n} fp{prec}: {r}", file=sys.stderr, flush=True)
        e64, e32 = extra["256^3_fp64"], extra["256^3_fp32"]
        extra["256^3_fp32_speedup_over_fp64"] = {
            "wall": e64["ms_per_step"] / e32["ms_per_step"],
            "events": e64["event_ms_per_step"] / e32["event_ms_per_step"]}

        r = run_conserve2(steps=xsteps)
        extra["conserve2_128^3_fp64"] = r
        print(f"# conserve 2: {r}", file=sys.stderr, flush=True)
        r = run_le(steps=xsteps)
        extra["lees_edwards_128^3_2planes_fp64"] = r
        print(f"# lees-edwards: {r}", file=sys.stderr, flush=True)
        r = run_ensemble(256, xsteps, 5)
        extra["ensemble_256_replicas_fp32"] = r
        print(f"# ensemble: {r}", file=sys.stderr, flush=True)
        extra["surrogate_gain"] = run_surrogate_gain()
        r = run_surrogate_conv()
        extra["surrogate_unet_inference_bf16_mfma"] = r
        print(f"# surrogate conv: {r}", file=sys.stderr, flush=True)

    roof = roofline(main_run, TRAFFIC_B_SITE.get(("binary_step", 64, args.size)))
    cpu = None if args.no_cpu else cpu_baseline()

    line = {
        "metric": "D3Q19 binary-fluid MLUPS & %HBM peak; surrogate speedup at <=2% L(t) error",
        "value": main_run["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_run["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded spinodal initial condition)",
        "config": {"workload": f"binary fluid D3Q19 MRT + symmetric FE + Cahn-Hilliard, "
                               f"{args.size}^3 periodic, FP64, tau=1, A=-0.0625, B=0.0625, "
                               f"kappa=0.04, M=0.25 (stated 0.45 is unstable), advection order 1",
                   "lattice": [args.size] * 3, "precision": "fp64"},
        "hbm_frac_step": main_run["hbm_frac_step"],
        "event_ms_per_step": main_run["event_ms_per_step"],
        "kernel_ms": main_run["kernel_ms"],
        "roofline": roof,
        "cpu_baseline": cpu,
        "surrogate": _surrogate_summary(extra.get("surrogate_gain")),
        "extra": extra,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
