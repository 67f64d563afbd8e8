"""The host driver (ludwig_amd/ludwig_hip) runs unchanged Ludwig input files
and its statistics lines are compared with logs written by the reference's
own Ludwig.exe (and two committed regression logs of the reference,
tests/regression/d3q19-short/serial-dist-3du.log and serial-spin-fd1.log,
kept here as .ref files with their unchanged .inp inputs; serial-spin-fd1
uses the 27-point gradient), token by token with the
absolute tolerance 1e-12 of the reference's tests/awk-fp-diff.sh."""

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


def run_driver(name, tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("ludwig_amd/ludwig_hip is not built")
    inp = os.path.join(LOGS, name + ".inp")
    out = subprocess.run([DRIVER, inp], capture_output=True, text=True, timeout=300,
                         cwd=tmp_path)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    return out.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["serial-dist-3du", "serial-spin-fd1", "spin7-o1", "spin7-o3",
                                  "c2-m025", "gradmu-trt", "single-bgk"])
def test_driver_matches_reference_log(name, tmp_path):
    got = run_driver(name, tmp_path)
    want = open(os.path.join(LOGS, name + ".ref")).read()
    print(name, "\n".join(stat_lines(got)[-4:]), flush=True)
    compare_logs(got, want)


def test_log_comparison_catches_a_change():
    """The comparison itself: a one-digit change in a printed value fails."""
    want = open(os.path.join(LOGS, "spin7-o1.ref")).read()
    lines = want.splitlines()
    i = next(k for k, l in enumerate(lines) if l.startswith("[phi]") and "e-04" in l)
    lines[i] = lines[i].replace("e-04", "e-03", 1)
    with pytest.raises(AssertionError):
        compare_logs("\n".join(lines), want)
    compare_logs(want, want)
