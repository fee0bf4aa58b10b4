"""Drop-in test: the reference tracer's own program (its objects compiled by
oracle/Makefile from the reference sources), once as built and once with
particle_tracking_hybrid replaced by test_particle_amd/host/tp_hybrid_shim.c
(the GPU drop-in, linked to libtpgpu.so).  Both run the same spec through
calc_energy_spectrum + particle_tracking_hybrid and must write the same
data/espectrum.dat and data/diff_coeffs.dat.

The binaries are built in the build container (`make -C oracle ref ref-gpu`)
and travel with the tree; without them the test is skipped."""
import os
import subprocess

import numpy as np
import pytest

from test_particle_amd.particles import init_particles

pytestmark = pytest.mark.gpu

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
REF = os.path.join(ROOT, "oracle", "_ref")
TWO_PI = 2.0 * np.pi


def _need(*names):
    for n in names:
        if not os.path.exists(os.path.join(REF, n)):
            pytest.skip(f"oracle/_ref/{n} not built")


def _spec(path, **kw):
    with open(path, "w") as f:
        for k, v in kw.items():
            f.write(f"{k} {float(v)!r}\n" if isinstance(v, float) else f"{k} {v}\n")


def _run_both(tmp, cpu_bin, gpu_bin, spec, ptl):
    outs = []
    for b in (cpu_bin, gpu_bin):
        d = os.path.join(tmp, b)
        os.makedirs(os.path.join(d, "data"), exist_ok=True)
        _spec(os.path.join(d, "spec.txt"), **spec)
        ptl.tofile(os.path.join(d, "p.bin"))
        print(f"[dropin] {b} {spec.get('system')}", flush=True)
        subprocess.run([os.path.join(REF, b), "hybrid", "spec.txt", "p.bin"], cwd=d,
                       check=True, timeout=120, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
        outs.append({f: open(os.path.join(d, "data", f)).read()
                     for f in sorted(os.listdir(os.path.join(d, "data")))})
    return outs


def test_dropin_pic_grid(tmp_path):
    """PIC grid (system 4): no sin/cos on the path, trajectories bitwise."""
    _need("ref_harness", "ref_harness_gpu")
    rng = np.random.default_rng(31)
    nz, ny, nx = 9, 8, 10
    lo, hi = (0.0, 0.0, 0.0), (9.0, 7.0, 8.0)
    g = rng.normal(0, 1, (2, nz, ny, nx, 3)).astype(np.float32)
    g[1, ..., 2] += 2.0
    g.tofile(os.path.join(tmp_path, "grid.bin"))
    ptl = init_particles(500, lo, hi, 0.3, seed=32)
    spec = dict(system=4, nx=nx, ny=ny, nz=nz, tmax=800 * 0.02 + 1e-9, dt=0.02, bc=0,
                nbins=60, nt_out=8, emin=1e-3, emax=1.0,
                grid=os.path.join(tmp_path, "grid.bin"),
                xmin=lo[0], ymin=lo[1], zmin=lo[2], xmax=hi[0], ymax=hi[1], zmax=hi[2])
    cpu, gpu = _run_both(str(tmp_path), "ref_harness", "ref_harness_gpu", spec, ptl)
    assert set(cpu) == {"diff_coeffs.dat", "espectrum.dat"}
    assert cpu == gpu


def test_dropin_open_boundary_no_escape(tmp_path):
    """System 0 with open boundaries through the reference's own writers
    (including data/espect_escape.dat).  No particle escapes here: once one
    does, the reference writes nptl_remain_time one entry past its end
    (src/tracking.c:366-369) and the reference program crashes in free();
    escapes are covered by the oracle tests instead."""
    _need("ref_harness", "ref_harness_gpu")
    lo, hi = (0.0, 0.0, 0.0), (5.0, 5.0, 5.0)
    ptl = init_particles(400, (2.0, 2.0, 2.0), (3.0, 3.0, 3.0), 0.1, seed=33)
    spec = dict(system=0, tmax=300 * 0.05 + 1e-9, dt=0.05, bc=1, nbins=40, nt_out=3,
                emin=1e-4, emax=1.0, xmin=lo[0], ymin=lo[1], zmin=lo[2], xmax=hi[0],
                ymax=hi[1], zmax=hi[2])
    cpu, gpu = _run_both(str(tmp_path), "ref_harness", "ref_harness_gpu", spec, ptl)
    assert "espect_escape.dat" in cpu
    assert cpu["espectrum.dat"] == gpu["espectrum.dat"]
    assert cpu["espect_escape.dat"] == gpu["espect_escape.dat"]
    # D_mumu/D_alphaalpha are rounding noise here (mu is conserved); compare
    # the other columns and the counts
    for lc, lg in zip(cpu["diff_coeffs.dat"].splitlines(), gpu["diff_coeffs.dat"].splitlines()):
        a, b = lc.split(), lg.split()
        assert a[:5] + a[7:] == b[:5] + b[7:]


def test_dropin_abc(tmp_path):
    """ABC field (system 2), 300 Wirz steps: within the 1e-10 horizon of the
    sin/cos difference, so the printed files agree."""
    _need("ref_harness_abc1", "ref_harness_gpu_abc1")
    dt = 0.01 * TWO_PI / np.sqrt(1 - 0.49) / np.sqrt(3.0)
    ptl = init_particles(256, (0, 0, 0), (TWO_PI,) * 3, 0.7, seed=34)
    spec = dict(system=2, nx=5, ny=5, nz=5, tmax=300 * dt + 1e-9, dt=dt, bc=0, nbins=102,
                nt_out=3, emin=1e-3, emax=10.0, xmin=0.0, ymin=0.0, zmin=0.0, xmax=TWO_PI,
                ymax=TWO_PI, zmax=TWO_PI)
    cpu, gpu = _run_both(str(tmp_path), "ref_harness_abc1", "ref_harness_gpu_abc1", spec, ptl)
    assert cpu["espectrum.dat"] == gpu["espectrum.dat"]
    for lc, lg in zip(cpu["diff_coeffs.dat"].splitlines()[1:],
                      gpu["diff_coeffs.dat"].splitlines()[1:]):
        a = np.array(lc.split(), dtype=float)
        b = np.array(lg.split(), dtype=float)
        assert np.allclose(a[:7], b[:7], rtol=2e-4) and a[8] == b[8]


def test_dropin_trajectory_diagnostics(tmp_path):
    """traj_diagnose = 1 through particle_tracking_hybrid: the reference
    program with the GPU shim fills ptl_time as the reference does (within
    1e-10: the GPU's sin/cos differ from libm by an ulp)."""
    _need("ref_harness", "ref_harness_gpu")
    rng = np.random.default_rng(41)
    nz, ny, nx = 7, 6, 8
    lo, hi = (0.0, 0.0, 0.0), (7.0, 5.0, 6.0)
    g = rng.normal(0, 1, (2, nz, ny, nx, 3)).astype(np.float32)
    g[1, ..., 2] += 2.0
    g.tofile(os.path.join(tmp_path, "grid.bin"))
    ptl = init_particles(100, lo, hi, 0.3, seed=42)
    steps, nout = 400, 9
    spec = dict(system=4, nx=nx, ny=ny, nz=nz, tmax=steps * 0.02 + 1e-9, dt=0.02, bc=0,
                nbins=40, nt_out=4, emin=1e-3, emax=1.0, nsteps_output=nout,
                grid=os.path.join(tmp_path, "grid.bin"),
                xmin=lo[0], ymin=lo[1], zmin=lo[2], xmax=hi[0], ymax=hi[1], zmax=hi[2])
    outs = []
    for b in ("ref_harness", "ref_harness_gpu"):
        d = os.path.join(str(tmp_path), b)
        os.makedirs(os.path.join(d, "o"), exist_ok=True)
        _spec(os.path.join(d, "spec.txt"), **spec)
        ptl.tofile(os.path.join(d, "p.bin"))
        print(f"[dropin] {b} trajh", flush=True)
        subprocess.run([os.path.join(REF, b), "trajh", "spec.txt", "p.bin", "o"], cwd=d,
                       check=True, timeout=120, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
        outs.append((np.fromfile(os.path.join(d, "o", "traj.bin")).reshape(100, -1, 11),
                     np.fromfile(os.path.join(d, "o", "ptl.bin")).reshape(-1, 11)))
    (tc, pc), (tg, pg) = outs
    assert tc.shape[1] == steps // nout + 1
    assert np.all(tc[:, -1, 7] > 0)            # the last slot was written
    scale = np.maximum(np.abs(tc), 1.0)
    assert np.max(np.abs(tg - tc) / scale) < 1e-10
    assert np.max(np.abs(pg - pc) / np.maximum(np.abs(pc), 1.0)) < 1e-10


def test_dropin_pic_grid_two_time_slices(tmp_path):
    """System 4 with two time slices (multi_tframe = 1): the reference
    program with the GPU shim writes the same espectrum.dat and
    diff_coeffs.dat as the reference."""
    _need("ref_harness", "ref_harness_gpu")
    rng = np.random.default_rng(51)
    nz, ny, nx = 9, 8, 10
    lo, hi = (0.0, 0.0, 0.0), (9.0, 7.0, 8.0)
    g = rng.normal(0, 1, (4, nz, ny, nx, 3)).astype(np.float32)   # E, B next; E, B prev
    g[1, ..., 2] += 2.0
    g[3, ..., 2] += 2.0
    g.tofile(os.path.join(tmp_path, "grid2.bin"))
    ptl = init_particles(400, lo, hi, 0.3, seed=52)
    spec = dict(system=4, nx=nx, ny=ny, nz=nz, tmax=600 * 0.02 + 1e-9, dt=0.02, bc=0,
                nbins=60, nt_out=6, emin=1e-3, emax=1.0, mt=1, grid_dt=1.3,
                grid=os.path.join(tmp_path, "grid2.bin"),
                xmin=lo[0], ymin=lo[1], zmin=lo[2], xmax=hi[0], ymax=hi[1], zmax=hi[2])
    cpu, gpu = _run_both(str(tmp_path), "ref_harness", "ref_harness_gpu", spec, ptl)
    assert set(cpu) == {"diff_coeffs.dat", "espectrum.dat"}
    assert cpu == gpu
