"""The host driver (ludwig_amd/ludwig_hip) runs unchanged Ludwig input files
and its statistics lines are compared with logs written by the reference's
own Ludwig.exe (and two committed regression logs of the reference,
tests/regression/d3q19-short/serial-dist-3du.log and serial-spin-fd1.log,
kept here as .ref files with their unchanged .inp inputs; serial-spin-fd1
uses the 27-point gradient), token by token with the
absolute tolerance 1e-12 of the reference's tests/awk-fp-diff.sh.

Lees-Edwards planes: the reference's regression inputs
tests/regression/d3q19-short/serial-le3d-st{1,...,8} (two planes of
steady shear on 32^3 from the shear-profile start, advection orders 1-4,
the 7- and 27-point gradients) and their committed logs.

External chemical potential gradient: serial-muex-st1 (grad_mu with the
phi_gradmu force, a droplet on 32 x 32 x 16, the 27-point gradient) and its
committed log.

Restart: the reference's tests/regression/d3q19-io/iodrop-mpi1-io{1,2,3}
run a droplet 0 -> 20 in one go, 0 -> 10 writing a configuration, and
10 -> 20 from that configuration.  The driver runs the same chain; its logs
match the reference's, the files it writes at step 10 are byte for byte
the ones Ludwig.exe writes (SHA-256 and metadata text recorded by
tests/golden/make_logs.py), and the restarted run ends exactly where the
single run ends.  The same with ascii records, and with two Lees-Edwards
planes (le-io1..3, from serial-le3d-st1: the metadata records the
planes)."""

import hashlib
import json
import os
import re
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LOGS = os.path.join(HERE, "golden", "logs")
DRIVER = os.path.join(REPO, "ludwig_amd", "ludwig_hip")

STAT = re.compile(r"^\[(rho|phi|fed|total   |fluid   |minimum |maximum )\]")


def stat_lines(text):
    out = []
    for line in text.splitlines():
        if STAT.match(line):
            out.append(line)
    return out


def tokens(line):
    tag, rest = line.split("]", 1)
    return tag + "]", [float(t) for t in re.findall(r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?|nan|inf", rest)]


def compare_logs(got, want, tol=1e-12):
    g, w = stat_lines(got), stat_lines(want)
    assert len(g) == len(w), (len(g), len(w), g, w)
    for a, b in zip(g, w):
        ta, va = tokens(a)
        tb, vb = tokens(b)
        assert ta == tb, (a, b)
        assert len(va) == len(vb), (a, b)
        for x, y in zip(va, vb):
            assert abs(x - y) <= tol, f"\n got  {a}\n want {b}"


def run_driver(name, tmp_path, env=None):
    if not os.path.exists(DRIVER):
        pytest.fail("ludwig_amd/ludwig_hip is not built")
    inp = os.path.join(LOGS, name + ".inp")
    out = subprocess.run([DRIVER, inp], capture_output=True, text=True, timeout=300,
                         cwd=tmp_path, env=None if env is None else dict(os.environ, **env))
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    return out.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["serial-dist-3du", "serial-spin-fd1", "spin7-o1", "spin7-o3",
                                  "c2-m025", "gradmu-trt", "single-bgk", "spin7-c1", "spin7-c2",
                                  "serial-le3d-st1", "serial-le3d-st2", "serial-le3d-st3",
                                  "serial-le3d-st5", "serial-le3d-st6", "serial-le3d-st7",
                                  "le-t0", "serial-muex-st1", "serial-le3d-st4", "serial-le3d-st8"])
def test_driver_matches_reference_log(name, tmp_path):
    got = run_driver(name, tmp_path)
    want = open(os.path.join(LOGS, name + ".ref")).read()
    print(name, "\n".join(stat_lines(got)[-4:]), flush=True)
    compare_logs(got, want)


@pytest.mark.gpu
def test_driver_conserve2_tree_sum_matches_log(tmp_path):
    """cahn_hilliard_options_conserve 2 with the library's default sum (the
    compensated tree reduction, not the one-thread site order): the
    statistics still match Ludwig.exe's log under the 1e-12 rule."""
    got = run_driver("spin7-c2", tmp_path, env={"LDW_CONSERVE2_SITE_ORDER": "0"})
    compare_logs(got, open(os.path.join(LOGS, "spin7-c2.ref")).read())


def test_log_comparison_catches_a_change():
    """The comparison itself: a one-digit change in a printed value fails."""
    want = open(os.path.join(LOGS, "spin7-o1.ref")).read()
    lines = want.splitlines()
    i = next(k for k, l in enumerate(lines) if l.startswith("[phi]") and "e-04" in l)
    lines[i] = lines[i].replace("e-04", "e-03", 1)
    with pytest.raises(AssertionError):
        compare_logs("\n".join(lines), want)
    compare_logs(want, want)


CHAINS = {"iodrop": ["iodrop-mpi1-io1", "iodrop-mpi1-io2", "iodrop-mpi1-io3"],
          "iodrop-ascii": ["iodrop-ascii-io2", "iodrop-ascii-io3"],
          "le-restart": ["le-io1", "le-io2", "le-io3"],
          # rho every 4 steps and velocity (ascii) every 6 between the
          # configuration steps, byte-identical with Ludwig.exe's files
          "hydro-io": ["hydro-io1"]}


@pytest.mark.gpu
@pytest.mark.parametrize("chain", list(CHAINS))
def test_restart_chain_matches_reference(chain, tmp_path):
    outs = {}
    for name in CHAINS[chain]:
        outs[name] = run_driver(name, tmp_path)
        want = open(os.path.join(LOGS, name + ".ref")).read()
        print(name, "\n".join(stat_lines(outs[name])[-4:]), flush=True)
        compare_logs(outs[name], want)
        if chain == "hydro-io":
            # which file was written at which step, in Ludwig.exe's order
            wr = lambda t: [l for l in t.splitlines() if l.startswith("Writing")]
            assert wr(outs[name]) == wr(want), (wr(outs[name]), wr(want))
    # the configuration files of step 10, against Ludwig.exe's
    ref = json.load(open(os.path.join(LOGS, chain + "-files.json")))
    for fn, d in sorted(ref.items()):
        path = tmp_path / fn
        assert path.exists(), f"{fn} not written"
        if "text" in d:
            assert path.read_text() == d["text"], fn
        else:
            data = path.read_bytes()
            assert len(data) == d["size"], fn
            assert hashlib.sha256(data).hexdigest() == d["sha256"], f"{fn} differs from the reference's"
    # 10 -> 20 from the files ends exactly where 0 -> 20 in one run ends
    # (with Lees-Edwards planes it does not, in the reference as here: the
    # plane offset uses t_start + t_current - 1, which counts N_start twice)
    if "iodrop-mpi1-io1" in outs:
        assert stat_lines(outs["iodrop-mpi1-io3"])[-7:] == stat_lines(outs["iodrop-mpi1-io1"])[-7:]


def test_reference_restart_is_exact_in_its_own_logs():
    """The golden logs themselves: Ludwig.exe's 10 -> 20 restart prints the
    same step-20 statistics as its 0 -> 20 run (momentum aside, whose last
    digits depend on the OpenMP summation order)."""
    one = [l for l in stat_lines(open(os.path.join(LOGS, "iodrop-mpi1-io1.ref")).read())[-7:]
           if not l.startswith(("[total", "[fluid"))]
    res = [l for l in stat_lines(open(os.path.join(LOGS, "iodrop-mpi1-io3.ref")).read())[-7:]
           if not l.startswith(("[total", "[fluid"))]
    assert one == res and len(one) == 5
    meta = json.load(open(os.path.join(LOGS, "iodrop-files.json")))
    d = json.loads(meta["dist-metadata.001-001"]["text"])
    assert d["io_element"]["Count"] == 19 and d["io_options"]["Record format"] == "binary"
    assert meta["dist-000000010.001-001"]["size"] == 32**3 * 19 * 8
    a = json.load(open(os.path.join(LOGS, "iodrop-ascii-files.json")))
    assert a["dist-000000010.001-001"]["size"] == 32**3 * 19 * 24
    assert a["vel-000000010.001-001"]["size"] == 32**3 * (3 * 23 + 1)
