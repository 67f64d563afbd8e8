"""Statistics logs and configuration-file digests written by the reference's
own Ludwig.exe, the golden data of tests/test_gpu_driver.py.

    make -C oracle ref                          # builds oracle/_ref/Ludwig.exe
    python tests/golden/make_logs.py            # every restart chain
    python tests/golden/make_logs.py spin7-o1   # single inputs by name

A single NAME runs logs/NAME.inp (as "input", which Ludwig.exe reads from
its working directory, main.c:38-49) in a scratch directory and keeps the
whole standard output as logs/NAME.ref.

A chain runs its inputs one after the other in one scratch directory, so a
restart input (N_start > 0) reads the files the previous input wrote at its
configuration step.  Each log is kept as logs/NAME.ref; the data files left
in the directory are recorded in logs/CHAIN-files.json as size and SHA-256,
and the metadata files as their text.
"""

import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
LOGS = os.path.join(HERE, "logs")
LUDWIG = os.path.join(REPO, "oracle", "_ref", "Ludwig.exe")

CHAINS = {
    # tests/regression/d3q19-io/iodrop-mpi1-io{1,2,3}.inp of the reference:
    # 0 -> 20 in one run; 0 -> 10 writing a configuration; 10 -> 20 from it
    "iodrop": ["iodrop-mpi1-io1", "iodrop-mpi1-io2", "iodrop-mpi1-io3"],
    # the same restart with ascii records
    "iodrop-ascii": ["iodrop-ascii-io2", "iodrop-ascii-io3"],
    # two Lees-Edwards planes (serial-le3d-st1 with advection order 3): the
    # metadata records the planes, and the restarted time counter
    "le-restart": ["le-io1", "le-io2", "le-io3"],
    # rho and velocity files between configuration steps (rho_io_freq,
    # vel_io_freq), the velocity in ascii records
    "hydro-io": ["hydro-io1"],
}


def run(name, cwd):
    shutil.copyfile(os.path.join(LOGS, name + ".inp"), os.path.join(cwd, "input"))
    out = subprocess.run([LUDWIG], cwd=cwd, capture_output=True, text=True, timeout=1800,
                         env=dict(os.environ, OMP_NUM_THREADS="1"))
    if out.returncode != 0:
        raise SystemExit(f"{name}: Ludwig.exe failed\n{out.stdout[-2000:]}{out.stderr[-2000:]}")
    with open(os.path.join(LOGS, name + ".ref"), "w") as f:
        f.write(out.stdout)
    print(f"{name}.ref written", flush=True)


def digests(cwd):
    files = {}
    for fn in sorted(os.listdir(cwd)):
        path = os.path.join(cwd, fn)
        if "-metadata." in fn:
            files[fn] = {"text": open(path).read()}
        elif fn.endswith(".001-001"):
            data = open(path, "rb").read()
            files[fn] = {"size": len(data), "sha256": hashlib.sha256(data).hexdigest()}
    return files


def main(argv):
    if not os.path.exists(LUDWIG):
        raise SystemExit("oracle/_ref/Ludwig.exe is not built (make -C oracle ref)")
    names = argv or list(CHAINS)
    for name in names:
        with tempfile.TemporaryDirectory() as cwd:
            if name in CHAINS:
                for step in CHAINS[name]:
                    run(step, cwd)
                with open(os.path.join(LOGS, name + "-files.json"), "w") as f:
                    json.dump(digests(cwd), f, indent=1, sort_keys=True)
                print(f"{name}-files.json written", flush=True)
            else:
                run(name, cwd)


if __name__ == "__main__":
    main(sys.argv[1:])
