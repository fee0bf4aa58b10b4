"""Drop-in test: the reference tracer's own program (its objects compiled by
oracle/Makefile from the reference sources), once as built and once with
particle_tracking_hybrid replaced by test_particle_amd/host/tp_hybrid_shim.c
(the GPU drop-in, linked to libtpgpu.so).  Both run the same spec through
calc_energy_spectrum + particle_tracking_hybrid and must write the same
data/espectrum.dat and data/diff_coeffs.dat.

The binaries are built in the build container (`make -C oracle ref ref-gpu`)
and travel with the tree; without them the test is skipped.  The drop-in runs
with TP_SHIM_REQUIRE_GPU=1, so a call the GPU path does not take aborts the
program instead of running the reference's CPU loop, and each test checks the
shim's stderr line that the GPU tracked; without libtpgpu.so the drop-in
binary does not start and the test fails."""
import os
import subprocess

import numpy as np
import pytest

from test_particle_amd import abi
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


GPU_MARK = "tp_gpu: particle_tracking_hybrid on the GPU"
PMAP_MARK = "tp_gpu: poincare_map on the GPU"


def _run(binary, args, cwd, mark=GPU_MARK, **env):
    """Runs an oracle/_ref binary; for a drop-in binary (name with _gpu) the
    GPU path is required and its stderr line checked."""
    gpu = "_gpu" in binary
    e = dict(os.environ, OMP_NUM_THREADS="1", **env)
    if gpu:
        e["TP_SHIM_REQUIRE_GPU"] = "1"
    r = subprocess.run([os.path.join(REF, binary), *args], cwd=cwd, timeout=120,
                       stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True, env=e)
    assert r.returncode == 0, f"{binary} {args[0]} exited {r.returncode}: {r.stderr[-2000:]}"
    if gpu:
        assert mark in r.stderr, f"{binary}: the GPU path did not run: {r.stderr[-2000:]}"
    return r


def _run_both(tmp, cpu_bin, gpu_bin, spec, ptl):
    outs = []
    for b in (cpu_bin, gpu_bin):
        d = os.path.join(tmp, b)
        os.makedirs(os.path.join(d, "data"), exist_ok=True)
        _spec(os.path.join(d, "spec.txt"), **spec)
        ptl.tofile(os.path.join(d, "p.bin"))
        print(f"[dropin] {b} {spec.get('system')}", flush=True)
        _run(b, ["hybrid", "spec.txt", "p.bin"], d)
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


GPU_SECTION = "------------------------------- GPU tracking ---------------------------------\n"


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dropin_gpu_section_boris(tmp_path, oracle_tp, precision):
    """init.dat's GPU tracking section picks the Boris pusher (and the FP32
    form) for the reference program's tracking loop: the drop-in's particles
    are, bit for bit, the library's run with those settings (FP64: also the
    oracle's Boris run), so the section reaches the kernels; the FP32 form's
    differ from the FP64 form's."""
    from test_particle_amd.gpu import Tracker
    _need("ref_harness_gpu")
    rng = np.random.default_rng(71)
    nz, ny, nx = 9, 8, 10
    lo, hi = (0.0, 0.0, 0.0), (9.0, 7.0, 8.0)
    g = rng.normal(0, 1, (2, nz, ny, nx, 3)).astype(np.float32)
    g[1, ..., 2] += 2.0
    g.tofile(os.path.join(tmp_path, "grid.bin"))
    ptl = init_particles(300, lo, hi, 0.3, seed=72)
    steps, dt = 400, 0.02
    spec = dict(system=4, nx=nx, ny=ny, nz=nz, tmax=steps * dt + 1e-9, dt=dt, bc=0,
                nbins=60, nt_out=4, emin=1e-3, emax=1.0,
                grid=os.path.join(tmp_path, "grid.bin"),
                xmin=lo[0], ymin=lo[1], zmin=lo[2], xmax=hi[0], ymax=hi[1], zmax=hi[2])
    d = os.path.join(str(tmp_path), "run")
    os.makedirs(os.path.join(d, "data"))
    _spec(os.path.join(d, "spec.txt"), **spec)
    ptl.tofile(os.path.join(d, "p.bin"))
    with open(os.path.join(d, "init.dat"), "w") as f:
        f.write("Run type (1: test particle simulation 2: Poincare map): 1\n" + GPU_SECTION +
                f"Pusher (wirz or boris): boris\nPrecision (fp64 or fp32): {precision}\n" +
                "-" * 78 + "\n")
    print(f"[dropin] ref_harness_gpu boris {precision}", flush=True)
    r = _run("ref_harness_gpu", ["hybrid", "spec.txt", "p.bin", "out.bin"], d)
    assert f"boris, {precision}, flags 0)" in r.stderr, r.stderr[-2000:]
    out = np.fromfile(os.path.join(d, "out.bin")).reshape(-1, 11)

    def params(prec):
        return abi.make_params(dt=dt, estep=steps, lo=lo, hi=hi, pusher=abi.TP_PUSH_BORIS,
                               precision=prec, nbins=60, nt_out=4, emin=1e-3, emax=1.0)

    tr = Tracker(0)
    try:
        tr.set_field_grid(g[0], g[1], lo, hi, abi.TP_GATHER_REF_COMPAT)
        runs = {}
        for name, prec in (("fp64", abi.TP_FP64), ("fp32", abi.TP_FP32_MIXED)):
            tr.load(ptl)
            tr.run(params(prec))
            runs[name] = tr.store()[0]
    finally:
        tr.close()
    assert np.array_equal(out, runs[precision])
    if precision == "fp64":
        ref = oracle_tp.track_fixed(oracle_tp.Field.grid(g[0], g[1], lo, hi), params(abi.TP_FP64),
                                    ptl)
        assert np.array_equal(out, ref["ptl"])
    else:
        assert not np.array_equal(out, runs["fp64"])
        assert np.median(np.abs(out[:, :3] - runs["fp64"][:, :3])) < 1e-3


def test_dropin_mhd_grid(tmp_path):
    """System 3 (MHD velocity and magnetic grids, E = B x v): the reference
    program with the GPU shim writes the reference's espectrum.dat and
    diff_coeffs.dat."""
    _need("ref_harness", "ref_harness_gpu")
    rng = np.random.default_rng(61)
    nz, ny, nx = 9, 8, 10
    lo, hi = (0.0, 0.0, 0.0), (9.0, 7.0, 8.0)
    v = rng.normal(0, 0.2, (nz, ny, nx, 3)).astype(np.float32)
    b = rng.normal(0, 1, (nz, ny, nx, 3)).astype(np.float32)
    b[..., 2] += 2.0
    np.concatenate([v.ravel(), b.ravel()]).tofile(os.path.join(tmp_path, "mgrid.bin"))
    ptl = init_particles(500, lo, hi, 0.3, seed=62)
    spec = dict(system=3, nx=nx, ny=ny, nz=nz, tmax=800 * 0.02 + 1e-9, dt=0.02, bc=0,
                nbins=60, nt_out=8, emin=1e-3, emax=1.0,
                grid=os.path.join(tmp_path, "mgrid.bin"),
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
        _run(b, ["trajh", "spec.txt", "p.bin", "o"], d)
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


def _pmap_run(tmp, binary, mode, spec, cfg):
    from test_particle_amd.poincare import write_config
    d = os.path.join(tmp, f"{binary}_{mode}")
    os.makedirs(os.path.join(d, "data"), exist_ok=True)
    _spec(os.path.join(d, "spec.txt"), **spec)
    write_config(os.path.join(d, "init.dat"), cfg)
    print(f"[dropin] {binary} {mode}", flush=True)
    _run(binary, [mode, "spec.txt", "init.dat", "out.bin"], d, mark=PMAP_MARK)
    np_ = 1 if cfg["single"] else cfg["nh"] * cfg["nv"]
    return np.fromfile(os.path.join(d, "out.bin")).reshape(3, np_, cfg["ncross"] + 1)


def _pmap_spec(tmp, name, e, b, lo, hi, **kw):
    path = os.path.join(tmp, f"{name}.bin")
    np.concatenate([a.ravel() for a in (*e, *b)]).astype(np.float32).tofile(path)
    nz, ny, nx = b[0].shape[:3]
    return dict(system=4, nx=nx, ny=ny, nz=nz, tmax=1.0, dt=0.01, grid=path,
                xmin=float(lo[0]), ymin=float(lo[1]), zmin=float(lo[2]), xmax=float(hi[0]),
                ymax=float(hi[1]), zmax=float(hi[2]), **kw)


@pytest.mark.parametrize("name", ["pmap_grid", "pmap_single", "pmap_grid2"])
def test_dropin_poincare_map_golden(tmp_path, name):
    """run_type 2 through the reference's poincare_map entry replaced by
    test_particle_amd/host/tp_pmap_shim.c: the reference's config parser
    reads the block, the lines are traced on the GPU and
    data/poincare_map.h5 holds the reference's crossing points bit for bit
    (read back from the file)."""
    import json
    from conftest import golden
    _need("ref_harness_gpu")
    g = golden(name)
    cfg = json.loads(str(g["cfg"]))
    if name == "pmap_grid2":
        z = np.zeros_like(g["b_next"])
        spec = _pmap_spec(str(tmp_path), name, (z, z), (g["b_next"], g["b_prev"]), g["lo"],
                          g["hi"], mt=1, grid_dt=float(g["grid_dt"]))
        # the file holds E, B of slice a then E, B of slice b
        np.concatenate([z.ravel(), g["b_next"].ravel(), z.ravel(),
                        g["b_prev"].ravel()]).tofile(spec["grid"])
    else:
        z = np.zeros_like(g["bgrid"])
        spec = _pmap_spec(str(tmp_path), name, (z,), (g["bgrid"],), g["lo"], g["hi"])
    out = _pmap_run(str(tmp_path), "ref_harness_gpu", "pmaph", spec, cfg)
    for k, key in enumerate(("xpos", "ypos", "zpos")):
        assert np.array_equal(out[k], g[key]), key


def test_dropin_poincare_map_turbulent_grid(tmp_path):
    """A 33^3 turbulent PIC grid, a 12 x 10 plane of start points: the
    reference (next_location driven by the harness) and the program with
    the GPU drop-in write the same map."""
    from test_particle_amd.fields import pic_turbulence_grid
    _need("ref_harness", "ref_harness_gpu")
    e, b, lo, hi = pic_turbulence_grid(n=32, seed=8, db_rms=0.8)
    spec = _pmap_spec(str(tmp_path), "turb", (np.zeros_like(b),), (b,), lo, hi)
    cfg = dict(face=2, step=0.05, length_max=150.0, ncross=8, single=0, x0=0.0, y0=0.0,
               z0=9.3, min_h=0.5, max_h=31.5, min_v=1.0, max_v=30.0, nh=12, nv=10)
    cpu = _pmap_run(str(tmp_path), "ref_harness", "pmap", spec, cfg)
    gpu = _pmap_run(str(tmp_path), "ref_harness_gpu", "pmaph", spec, cfg)
    assert np.array_equal(cpu, gpu)
    filled = (cpu[0] != 2 * hi[0]).sum(axis=1) - 1
    assert filled.max() == 8 and filled.min() < 8
