"""The drop-in shim's failure handling, on a machine without a GPU.

oracle/_ref/ref_harness_gpu is the reference program with
particle_tracking_hybrid replaced by test_particle_amd/host/tp_hybrid_shim.c.
Without a GPU, tp_gpu_init fails: the shim says so on stderr and runs the
reference's own CPU loop (the files match the reference program's), and with
TP_SHIM_REQUIRE_GPU=1 it aborts instead.  Skipped where a GPU is visible
(/dev/kfd), since there the GPU path runs."""
import os
import subprocess

import numpy as np
import pytest

from test_particle_amd.particles import init_particles

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
REF = os.path.join(ROOT, "oracle", "_ref")

pytestmark = pytest.mark.skipif(
    os.path.exists("/dev/kfd") or not os.path.exists(os.path.join(REF, "ref_harness_gpu")),
    reason="needs the drop-in binary and no visible GPU")


def _spec(d):
    lo, hi = (0.0, 0.0, 0.0), (5.0, 5.0, 5.0)
    ptl = init_particles(50, lo, hi, 0.1, seed=3)
    ptl.tofile(os.path.join(d, "p.bin"))
    spec = dict(system=0, tmax=200 * 0.05 + 1e-9, dt=0.05, bc=0, nbins=20, nt_out=2,
                emin=1e-4, emax=1.0, xmin=0.0, ymin=0.0, zmin=0.0, xmax=5.0, ymax=5.0,
                zmax=5.0)
    with open(os.path.join(d, "spec.txt"), "w") as f:
        for k, v in spec.items():
            f.write(f"{k} {v!r}\n" if isinstance(v, float) else f"{k} {v}\n")


def _run(binary, d, init=None, **env):
    os.makedirs(os.path.join(d, "data"), exist_ok=True)
    _spec(d)
    if init is not None:
        with open(os.path.join(d, "init.dat"), "w") as f:
            f.write(init)
    return subprocess.run([os.path.join(REF, binary), "hybrid", "spec.txt", "p.bin"], cwd=d,
                          timeout=120, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                          text=True, env=dict(os.environ, OMP_NUM_THREADS="1", **env))


def test_shim_falls_back_visibly_without_gpu(tmp_path):
    ref = _run("ref_harness", str(tmp_path / "cpu"))
    shim = _run("ref_harness_gpu", str(tmp_path / "shim"))
    assert ref.returncode == 0 and shim.returncode == 0
    assert "particle_tracking_hybrid runs the CPU loop: no usable GPU" in shim.stderr
    assert "on the GPU" not in shim.stderr
    for f in ("espectrum.dat", "diff_coeffs.dat"):
        a = open(tmp_path / "cpu" / "data" / f).read()
        b = open(tmp_path / "shim" / "data" / f).read()
        assert a == b, f


def test_shim_aborts_when_gpu_required(tmp_path):
    r = _run("ref_harness_gpu", str(tmp_path / "shim"), TP_SHIM_REQUIRE_GPU="1")
    assert r.returncode != 0
    assert "TP_SHIM_REQUIRE_GPU is set; aborting" in r.stderr


SECTION = "------------------------------- GPU tracking ---------------------------------\n"
RULE = "-" * 78 + "\n"


def test_shim_gpu_section_with_the_reference_settings_falls_back(tmp_path):
    """init.dat's GPU tracking section asking for what the reference's loop
    does (Wirz, FP64): without a GPU the CPU loop runs as before."""
    init = (SECTION + "Pusher (wirz or boris): wirz\nPrecision (fp64 or fp32): fp64\n"
            "Adaptive relative tolerance: 1e-12\n" + RULE)
    ref = _run("ref_harness", str(tmp_path / "cpu"))
    shim = _run("ref_harness_gpu", str(tmp_path / "shim"), init=init)
    assert ref.returncode == 0 and shim.returncode == 0, shim.stderr[-2000:]
    assert "runs the CPU loop: no usable GPU" in shim.stderr
    for f in ("espectrum.dat", "diff_coeffs.dat"):
        assert (open(tmp_path / "cpu" / "data" / f).read() ==
                open(tmp_path / "shim" / "data" / f).read()), f


def test_shim_gpu_section_boris_does_not_fall_back(tmp_path):
    """The reference's loop has no Boris pusher: asked for one, the drop-in
    without a GPU stops instead of running the Wirz loop."""
    r = _run("ref_harness_gpu", str(tmp_path / "shim"),
             init=SECTION + "Pusher (wirz or boris): boris\n" + RULE)
    assert r.returncode != 0
    assert "asks for what the CPU loop does not do (pusher boris" in r.stderr
    assert not os.path.exists(tmp_path / "shim" / "data" / "diff_coeffs.dat")


def test_shim_gpu_section_malformed_stops(tmp_path):
    """A malformed GPU tracking section stops the program with the line and
    the reason, as the reference's own parser stops on a bad line
    (src/domain.c:93-96)."""
    r = _run("ref_harness_gpu", str(tmp_path / "shim"),
             init="Run type: 1\n" + SECTION + "Precision (fp64 or fp32): fp16\n" + RULE)
    assert r.returncode != 0
    assert 'tp_gpu: init.dat line 3: precision "fp16" is neither fp64 nor fp32' in r.stderr
    # the CPU loop did not run: it would have written the diffusion file
    assert not os.path.exists(tmp_path / "shim" / "data" / "diff_coeffs.dat")
