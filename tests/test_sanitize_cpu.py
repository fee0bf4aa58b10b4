"""The CPU code under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle sanitize` builds oracle/sanitize_check.c with the oracle
(every field kind, both pushers, the boundary conditions, the adaptive
stepper with and without dense output, the Poincare map, the surrogate
network), the product's host C (spectrum files, .gda field input, Poincare
start points) and the drop-in's init.dat "GPU tracking" section reader,
all with -fsanitize=address,undefined -fno-sanitize-recover=all.  Any
sanitizer report ends the program with a non-zero status.

The section reader's results on the files given here are checked line by
line: the values it takes, its defaults (the reference's Wirz pusher in
FP64 and its Bulirsch-Stoer constants, include/StepperBS.h:18-22) and its
error messages."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
EXE = os.path.join(ROOT, "oracle", "_san", "sanitize_check")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None and shutil.which("cc") is None,
                                reason="needs gcc")

RULE = "-" * 78 + "\n"
SECTION = "------------------------------- GPU tracking ---------------------------------\n"

CONFIGS = {
    # the reference's own layout around the section; every line set
    "full": (
        "---------------------------- Simulation run type -----------------------------\n"
        "Run type (1: test particle simulation 2: Poincare map): 1\n" + RULE + "\n" +
        SECTION +
        "Pusher (wirz or boris): boris\n"
        "Precision (fp64 or fp32): fp32\n"
        "Check for non-finite states (0 or 1): 1\n"
        "Fix stale diffusion terms of escaped particles (0 or 1): 1\n"
        "\n"
        "Adaptive relative tolerance: 1e-10\n"
        "Adaptive absolute tolerance: 2.5e-8\n"
        "Adaptive first step: 0.001\n"
        "Adaptive maximum sequence index: 6\n" + RULE +
        "Tracking method to use: 0\n"),
    # no section: the defaults
    "absent": "Run type (1: test particle simulation 2: Poincare map): 1\n" + RULE,
    # a section that runs to the end of the file, CRLF line ends, a flag set
    # and cleared again
    "eof": SECTION + "Pusher: wirz\r\nCheck for non-finite states: 1\r\n"
                     "Check for non-finite states: 0\r\n",
    # a long line before the section is skipped, like any other line
    "longline": "x" * 3000 + "\n" + SECTION + "Precision (fp64 or fp32): fp64\n" + RULE,
    "badpusher": SECTION + "Pusher (wirz or boris): rk4\n" + RULE,
    "badlabel": SECTION + "Pusher (wirz or boris): boris\nTile size: 64\n" + RULE,
    "novalue": SECTION + "Pusher boris\n" + RULE,
    "fp32wirz": SECTION + "Precision (fp64 or fp32): fp32\n" + RULE,
    "badrtol": SECTION + "Adaptive relative tolerance: 2.0\n" + RULE,
    "badkmax": SECTION + "Adaptive maximum sequence index: 9\n" + RULE,
    "badflag": SECTION + "Check for non-finite states (0 or 1): yes\n" + RULE,
    "trailing": SECTION + "Adaptive first step: 1e-3x\n" + RULE,
    "longinside": SECTION + "Pusher: " + "b" * 3000 + "\n" + RULE,
}

DEFAULTS = dict(found=0, pusher=0, precision=0, flags=0, rtol=1e-14, atol=1e-5, h0=1e-5,
                kmax=8)


def _parse(line):
    # "config <path>: found 1 pusher 1 ... kmax 6" or "config <path>: error: <msg>"
    path, rest = line[len("config "):].split(": ", 1)
    if rest.startswith("error: "):
        return os.path.basename(path), rest[len("error: "):]
    w = rest.split()
    vals = {w[i]: float(w[i + 1]) for i in range(0, len(w), 2)}
    return os.path.basename(path), vals


@pytest.fixture(scope="module")
def sanitized():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return EXE


def test_cpu_code_is_sanitizer_clean(sanitized, tmp_path):
    paths = []
    for name, text in CONFIGS.items():
        p = tmp_path / name
        p.write_bytes(text.encode())
        paths.append(str(p))
    paths.append(str(tmp_path / "missing"))
    r = subprocess.run([sanitized, *paths], capture_output=True, text=True, timeout=300,
                       cwd=tmp_path, env=dict(os.environ, OMP_NUM_THREADS="2"))
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    lines = r.stdout.splitlines()
    assert lines[-1] == "sanitize_check: ok"
    parts = {ln.split()[0] + " " + ln.split()[1] for ln in lines if ln.startswith(
        ("fixed", "eval", "bs", "pmap"))}
    for kind in ("test", "grid", "grid2", "grid-tb", "mhd", "grid-f64", "abc", "abc-v", "modes",
                 "wlcs", "wlcs-j"):
        assert f"fixed {kind}" in parts and f"eval {kind}" in parts, kind
    assert {"bs test", "bs grid", "bs modes", "pmap grid"} <= parts
    assert any(ln.startswith("mlp ") for ln in lines)
    assert "host  spectrum files and .gda input ok" in lines

    got = dict(_parse(ln) for ln in lines if ln.startswith("config "))
    assert got["full"] == dict(found=1, pusher=1, precision=1, flags=3, rtol=1e-10,
                               atol=2.5e-8, h0=1e-3, kmax=6)
    assert got["absent"] == DEFAULTS
    assert got["missing"] == DEFAULTS
    assert got["eof"] == dict(DEFAULTS, found=1)
    assert got["longline"] == dict(DEFAULTS, found=1)
    assert got["badpusher"] == 'init.dat line 2: pusher "rk4" is neither wirz nor boris'
    assert got["badlabel"] == "init.dat line 3: unknown line in the GPU tracking section"
    assert got["novalue"] == ('init.dat line 2: expected "Label: value" in the GPU tracking '
                              'section')
    assert got["fp32wirz"].startswith("init.dat: precision fp32 needs pusher boris")
    assert got["badrtol"] == 'init.dat line 2: relative tolerance "2.0" is not in [0, 1]'
    assert got["badkmax"] == 'init.dat line 2: maximum sequence index "9" is not in 3..8'
    assert got["badflag"] == "init.dat line 2: expected 0 or 1"
    assert got["trailing"].startswith('init.dat line 2: first step "1e-3x"')
    assert got["longinside"].startswith("init.dat line 2: line longer than")
