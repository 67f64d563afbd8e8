#!/usr/bin/env python3
"""Write the golden vectors under tests/golden/ from the reference's own build.

Run in the development container only (it needs the reference sources):

    make -C oracle ref          # compiles the reference's sources into oracle/_ref/
    python tests/golden/make_golden.py

``oracle/_ref/ref_harness`` (source: ``oracle/ref_harness.c``) calls the
reference's public API -- lb_collide, lb_propagation, field_grad_compute,
phi_force_calculation, phi_cahn_hilliard, field_phi_init_spinodal -- on
seeded inputs and writes raw float64 arrays.  This script packs each case
into ``<case>.npz``: inputs, outputs and the parameters.  Large arrays are
stored as a SHA-256 of their bytes plus values at sampled sites, which is
enough to check bit-exactness and a relative tolerance.

The fixtures are data (inputs and expected outputs); no reference source
is copied.
"""

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")

# C2 parameters (symmetric free energy, Cahn-Hilliard)
A, B, K, M = -0.0625, 0.0625, 0.04, 0.45
MS = 0.25   # largest round mobility that is stable for |phi| <= 1.1 with A, B, K

# name: (harness args, shape, arrays stored in full, arrays stored by hash+samples)
CASES = {
    "collide_m10": (["collide", 8, 0, 1, 0.8], 8, ["f_in", "force", "f_out", "rho", "u"], []),
    "collide_bgk": (["collide", 8, 1, 2, 0.8], 8, ["f_in", "force", "f_out", "rho", "u"], []),
    "collide_trt": (["collide", 8, 2, 3, 0.8], 8, ["f_in", "force", "f_out", "rho", "u"], []),
    "collide_m10_tau1": (["collide", 8, 0, 4, 1.0], 8, ["f_in", "force", "f_out", "rho", "u"], []),
    "prop": (["prop", 8, 5], 8, ["f_in", "f_out"], []),
    "force_stress": (["force", 8, 6, 0, A, B, K], 8, ["phi", "grad", "delsq", "mu", "force"], []),
    "force_gradmu": (["force", 8, 7, 1, A, B, K], 8, ["phi", "grad", "delsq", "mu", "force"], []),
    "ch_order1": (["ch", 8, 8, 1, A, B, K, M], 8, ["phi_in", "u", "phi_out"], []),
    "ch_order2": (["ch", 8, 9, 2, A, B, K, M], 8, ["phi_in", "u", "phi_out"], []),
    "ch_order3": (["ch", 8, 10, 3, A, B, K, M], 8, ["phi_in", "u", "phi_out"], []),
    # full binary step, C2 physics (tau = 1), spinodal start
    "binary16_o1_s20": (["binary", 16, 20, 8361235, 1, A, B, K, M, 1.0, 0.0, 0.1, 10], 16,
                        ["phi_init", "phi_10", "phi", "rho", "u"], ["f"]),
    "binary16_o3_s10": (["binary", 16, 10, 4242, 3, A, B, K, M, 1.0, 0.0, 0.1], 16,
                        ["phi", "rho", "u"], ["f"]),
    "binary16_o2_s5": (["binary", 16, 5, 77, 2, A, B, K, M, 0.8, 0.1, 0.2], 16,
                       ["phi", "rho", "u"], ["f"]),
    # C2 physics with M = 0.25: the stated M = 0.45 is linearly unstable for
    # the explicit Cahn-Hilliard step (M k^2 (A + 3B phi^2 + kappa k^2) > 2 at
    # the grid scale) and the reference itself reaches NaN by step ~30
    "binary32_o1_s100": (["binary", 32, 100, 2024, 1, A, B, K, MS, 1.0, 0.0, 0.1], 32,
                         [], ["phi", "rho", "u", "f"]),
    "binary32_o3_s200": (["binary", 32, 200, 31337, 3, A, B, K, MS, 1.0, 0.0, 0.1], 32,
                         [], ["phi", "rho", "u", "f"]),
    # the reference's own run at the stated M = 0.45: finite at step 20,
    # non-finite by step 40
    "binary32_m045_s20": (["binary", 32, 20, 2024, 1, A, B, K, M, 1.0, 0.0, 0.1], 32,
                          [], ["phi", "rho", "u", "f"]),
    "binary32_m045_s40": (["binary", 32, 40, 2024, 1, A, B, K, M, 1.0, 0.0, 0.1], 32,
                          [], ["phi"]),
    # C1: plain D3Q19 MRT, 32^3, tau 0.8, 100 steps, random-velocity start
    "single32_s100": (["single", 32, 100, 12345, 0.8], 32, [], ["u_init", "f_init", "f", "rho", "u"]),
    # the 27-point gradient (fd_gradient_calculation 3d_27pt_fluid)
    "force27_stress": (["force", 8, 11, 0, A, B, K], 8, ["phi", "grad", "delsq", "mu", "force"], [],
                       {"LDW_HARNESS_STENCIL": "27"}),
    "force27_gradmu": (["force", 8, 12, 1, A, B, K], 8, ["phi", "grad", "delsq", "mu", "force"], [],
                       {"LDW_HARNESS_STENCIL": "27"}),
    # serial-spin-fd1.inp physics (A = -0.00625, B = 0.00625, K = 0.004,
    # M = 1.25, viscosity 0.00625 -> tau = 0.51875) with the 27-point gradient
    "binary16_g27_s10": (["binary", 16, 10, 8361235, 1, -0.00625, 0.00625, 0.004, 1.25, 0.51875, 0.0, 0.1],
                         16, ["phi", "rho", "u"], ["f"], {"LDW_HARNESS_STENCIL": "27"}),
    "binary16_g27_o3_s10": (["binary", 16, 10, 99, 3, A, B, K, MS, 1.0, 0.0, 0.1],
                            16, ["phi", "rho", "u"], ["f"], {"LDW_HARNESS_STENCIL": "27"}),
    # cahn_hilliard_options_conserve 1 (per-site compensated update) and 2
    # (global subtraction; its sum depends on the OpenMP thread count, so
    # the reference runs on one thread, the order the oracle and the GPU
    # library reproduce)
    "binary16_c1_s20": (["binary", 16, 20, 8361235, 1, A, B, K, MS, 1.0, 0.05, 0.1],
                        16, ["phi", "rho", "u", "csum"], ["f"],
                        {"LDW_HARNESS_CONSERVE": "1", "OMP_NUM_THREADS": "1"}),
    "binary16_c1_o3_s10": (["binary", 16, 10, 515, 3, A, B, K, MS, 0.8, -0.1, 0.2],
                           16, ["phi", "rho", "u", "csum"], ["f"],
                           {"LDW_HARNESS_CONSERVE": "1", "OMP_NUM_THREADS": "1"}),
    "binary16_c2_s20": (["binary", 16, 20, 8361235, 1, A, B, K, MS, 1.0, 0.05, 0.1],
                        16, ["phi_init_sum", "phi", "rho", "u"], ["f"],
                        {"LDW_HARNESS_CONSERVE": "2", "OMP_NUM_THREADS": "1"}),
    "binary16_c2_o2_s10": (["binary", 16, 10, 616, 2, A, B, K, MS, 0.8, -0.1, 0.2],
                           16, ["phi_init_sum", "phi", "rho", "u"], ["f"],
                           {"LDW_HARNESS_CONSERVE": "2", "OMP_NUM_THREADS": "1"}),
    # Lees-Edwards planes (N_LE_plane 2, steady shear): single calls at
    # time step 9 (buffer displacement 8 uy), the LB plane conditions, and
    # the whole step loop with the planes, as ludwig.c runs it
    "le_force16": (["force", 16, 21, 0, A, B, K], 16, ["phi", "grad", "delsq", "mu", "force"], [],
                   {"LDW_HARNESS_LE": "2 0.37", "LDW_HARNESS_T0": "9"}),
    "le_force16_g27": (["force", 16, 22, 0, A, B, K], 16, ["phi", "grad", "delsq", "mu", "force"], [],
                       {"LDW_HARNESS_LE": "2 0.37", "LDW_HARNESS_T0": "9", "LDW_HARNESS_STENCIL": "27"}),
    "le_ch16_o1": (["ch", 16, 23, 1, A, B, K, M], 16, ["phi_in", "u", "phi_out"], [],
                   {"LDW_HARNESS_LE": "2 0.37", "LDW_HARNESS_T0": "9"}),
    "le_ch16_o2": (["ch", 16, 24, 2, A, B, K, M], 16, ["phi_in", "u", "phi_out"], [],
                   {"LDW_HARNESS_LE": "2 0.37", "LDW_HARNESS_T0": "9"}),
    "le_ch16_o3": (["ch", 16, 25, 3, A, B, K, M], 16, ["phi_in", "u", "phi_out"], [],
                   {"LDW_HARNESS_LE": "2 0.37", "LDW_HARNESS_T0": "9"}),
    "le_lbbc16": (["lebc", 16, 26], 16, ["f_in", "f_out"], [],
                  {"LDW_HARNESS_LE": "2 0.37", "LDW_HARNESS_T0": "9"}),
    "le_binary16_o1_s20": (["binary", 16, 20, 8361235, 1, A, B, K, MS, 0.8, 0.0, 0.1], 16,
                           ["phi", "rho", "u"], ["f"],
                           {"LDW_HARNESS_LE": "2 0.05", "LDW_HARNESS_LE_PROFILE": "1"}),
    "le_binary16_o3_g27_s10": (["binary", 16, 10, 4242, 3, A, B, K, MS, 1.0, 0.0, 0.1], 16,
                               ["phi", "rho", "u"], ["f"],
                               {"LDW_HARNESS_LE": "2 0.1", "LDW_HARNESS_STENCIL": "27"}),
    "le_binary24_p3_o2_s12": (["binary", 24, 12, 77, 2, A, B, K, MS, 0.8, 0.1, 0.2], 24,
                              ["phi", "rho", "u"], ["f"],
                              {"LDW_HARNESS_LE": "3 0.2", "LDW_HARNESS_LE_PROFILE": "1"}),
    "le_binary16_c1_s10": (["binary", 16, 10, 515, 1, A, B, K, MS, 1.0, 0.05, 0.1], 16,
                           ["phi", "rho", "u", "csum"], ["f"],
                           {"LDW_HARNESS_LE": "2 0.05", "LDW_HARNESS_CONSERVE": "1",
                            "OMP_NUM_THREADS": "1"}),
    # LE_time_offset 7: the buffers move with t - 7, the Cahn-Hilliard faces with t
    "le_binary16_t0_s12": (["binary", 16, 12, 99, 2, A, B, K, MS, 0.8, 0.0, 0.1], 16,
                           ["phi", "rho", "u"], ["f"],
                           {"LDW_HARNESS_LE": "2 0.15 7", "LDW_HARNESS_LE_PROFILE": "1"}),
    # external chemical potential gradient (grad_mu): the Cahn-Hilliard face
    # fluxes and, with phi_gradmu, the force
    "force_gradmu_ext": (["force", 8, 31, 1, A, B, K], 8, ["phi", "grad", "delsq", "mu", "force"], [],
                         {"LDW_HARNESS_GRADMU": "0.001 -0.002 0.003"}),
    "ch_gradmu_o2": (["ch", 8, 32, 2, A, B, K, M], 8, ["phi_in", "u", "phi_out"], [],
                     {"LDW_HARNESS_GRADMU": "0.001 -0.002 0.003"}),
    "binary16_gmu_s10": (["binary", 16, 10, 8361235, 1, -0.00625, 0.00625, 0.004, 1.25, 0.51875, 0.0, 0.1],
                         16, ["phi", "rho", "u"], ["f"],
                         {"LDW_HARNESS_STENCIL": "27", "LDW_HARNESS_METHOD": "1",
                          "LDW_HARNESS_GRADMU": "0.00001 0.00002 0.00003"}),
    "binary16_gmu_stress_s10": (["binary", 16, 10, 5151, 2, A, B, K, MS, 1.0, 0.0, 0.1],
                                16, ["phi", "rho", "u"], ["f"],
                                {"LDW_HARNESS_GRADMU": "0.0005 0.0 -0.001"}),
    # advection orders 4 and 5 (order 5 with a coordinate halo of 3: its
    # stencil reaches three sites)
    "ch_order4": (["ch", 8, 33, 4, A, B, K, M], 8, ["phi_in", "u", "phi_out"], []),
    "ch_order5": (["ch", 8, 34, 5, A, B, K, M], 8, ["phi_in", "u", "phi_out"], [],
                  {"LDW_HARNESS_NHALO": "3"}),
    "binary16_o4_s10": (["binary", 16, 10, 4343, 4, A, B, K, MS, 1.0, 0.0, 0.1], 16,
                        ["phi", "rho", "u"], ["f"]),
    "binary16_o5_s10": (["binary", 16, 10, 5454, 5, A, B, K, MS, 1.0, 0.0, 0.1], 16,
                        ["phi", "rho", "u"], ["f"], {"LDW_HARNESS_NHALO": "3"}),
    "binary16_o4_g27_c1_s10": (["binary", 16, 10, 77, 4, A, B, K, MS, 0.8, 0.05, 0.2], 16,
                               ["phi", "rho", "u", "csum"], ["f"],
                               {"LDW_HARNESS_STENCIL": "27", "LDW_HARNESS_CONSERVE": "1",
                                "OMP_NUM_THREADS": "1"}),
    "le_ch16_o4": (["ch", 16, 35, 4, A, B, K, M], 16, ["phi_in", "u", "phi_out"], [],
                   {"LDW_HARNESS_LE": "2 0.37", "LDW_HARNESS_T0": "9"}),
    "le_binary16_o4_s10": (["binary", 16, 10, 13, 4, A, B, K, MS, 0.8, 0.0, 0.1], 16,
                           ["phi", "rho", "u"], ["f"],
                           {"LDW_HARNESS_LE": "2 0.05", "LDW_HARNESS_LE_PROFILE": "1"}),
    "le_binary16_gmu_s10": (["binary", 16, 10, 21, 2, A, B, K, MS, 0.8, 0.0, 0.1], 16,
                            ["phi", "rho", "u"], ["f"],
                            {"LDW_HARNESS_LE": "2 0.05", "LDW_HARNESS_LE_PROFILE": "1",
                             "LDW_HARNESS_GRADMU": "0.0005 0.0 -0.001"}),
}

NSAMPLE = 256


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def main():
    """python make_golden.py [case ...]: all cases, or only the named ones
    (the sample sites of every case stay the same either way)."""
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: make -C oracle ref")
    only = set(sys.argv[1:])
    rng = np.random.default_rng(20261202)
    for name, spec in CASES.items():
        args, n, full, hashed = spec[:4]
        env = dict(os.environ, OMP_NUM_THREADS="8")
        env.update(spec[4] if len(spec) > 4 else {})
        nsite = n ** 3
        samples = np.sort(rng.choice(nsite, size=min(NSAMPLE, nsite), replace=False))
        if only and name not in only:
            continue
        with tempfile.TemporaryDirectory() as tmp:
            cmd = [HARNESS, tmp] + [str(a) for a in args]
            subprocess.run(cmd, check=True, env=env)
            out = {"args": np.array([str(a) for a in args]), "n": np.array(n)}
            if len(spec) > 4:
                out["stencil"] = np.array(int(spec[4].get("LDW_HARNESS_STENCIL", 7)))
                out["conserve"] = np.array(int(spec[4].get("LDW_HARNESS_CONSERVE", 0)))
                if "LDW_HARNESS_GRADMU" in spec[4]:
                    out["grad_mu"] = np.array([float(v) for v in spec[4]["LDW_HARNESS_GRADMU"].split()])
                if "LDW_HARNESS_METHOD" in spec[4]:
                    out["force_method"] = np.array(int(spec[4]["LDW_HARNESS_METHOD"]))
                if "LDW_HARNESS_NHALO" in spec[4]:
                    out["nhalo"] = np.array(int(spec[4]["LDW_HARNESS_NHALO"]))
                if "LDW_HARNESS_LE" in spec[4]:
                    le = spec[4]["LDW_HARNESS_LE"].split()
                    out["le_planes"] = np.array(int(le[0]))
                    out["le_uy"] = np.array(float(le[1]))
                    out["le_nt0"] = np.array(int(le[2]) if len(le) > 2 else 0)
                    out["le_t0"] = np.array(int(spec[4].get("LDW_HARNESS_T0", 0)))
                    out["le_profile"] = np.array(int(spec[4].get("LDW_HARNESS_LE_PROFILE", 0)))
            out["sample_sites"] = samples
            for key in full:
                out[key] = np.fromfile(os.path.join(tmp, key + ".f64"))
            for key in hashed:
                a = np.fromfile(os.path.join(tmp, key + ".f64"))
                ncomp = a.size // nsite
                out[key + "_sha256"] = np.array(sha(a))
                out[key + "_samples"] = a.reshape(ncomp, nsite)[:, samples]
                out[key + "_sum"] = np.array(a.reshape(ncomp, nsite).sum(axis=1))
                out[key + "_finite"] = np.array(bool(np.isfinite(a).all()))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print("wrote", name)


if __name__ == "__main__":
    main()
