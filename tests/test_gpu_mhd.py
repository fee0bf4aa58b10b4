"""System 3 (MHD velocity and magnetic grids, E = B x v,
getemf_mhd_test_particle, src/emfields.c:107-120) and PIC grids of doubles
(get_double_efield_at_point / get_double_bfield_at_point,
src/electric_field.c:370-406, src/magnetic_field.c:555-591) on the GPU.

Fields bitwise the reference binary's golden vectors and the oracle (no
sin/cos on these paths); trajectories, spectra and counts of the fixed-step
loop bitwise the oracle (Wirz and Boris, cell records on and off), D sums
to 1e-12; the FP32 form of system 3 within float rounding of FP64; the
adaptive stepper and the Poincare map on the MHD field against the same on
the equivalent PIC grid."""
import json

import numpy as np
import pytest

from conftest import golden
from test_particle_amd import abi
from test_particle_amd.particles import init_particles

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from test_particle_amd.gpu import Tracker
    t = Tracker(0)
    yield t
    t.close()


def bitwise(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    assert not bad.any(), f"{what}: {bad.sum()} differ, first {np.argwhere(bad)[0]}"


def _pts(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(lo) - 0.3, np.asarray(hi) + 0.3
    return np.column_stack([rng.uniform(lo[k], hi[k], n) for k in range(3)] + [np.zeros(n)])


@pytest.mark.parametrize("cells", [abi.TP_CELLS_OFF, abi.TP_CELLS_ON])
def test_mhd_field(gpu, oracle_tp, cells):
    g = golden("fields_mhd")
    gpu.set_grid_cells(cells)
    try:
        gpu.set_field_mhd(g["vgrid"], g["bgrid"], g["lo"], g["hi"])
    finally:
        gpu.set_grid_cells(abi.TP_CELLS_AUTO)
    print(f"[mhd] field, cells mode {cells}", flush=True)
    bitwise(gpu.eval_field(g["pts"]), g["out"], "mhd vs reference")
    pts = _pts(2000, g["lo"], g["hi"], 5)
    f = oracle_tp.Field.mhd(g["vgrid"], g["bgrid"], g["lo"], g["hi"])
    bitwise(gpu.eval_field(pts), oracle_tp.eval_field(f, pts), "mhd vs oracle")


def test_double_grid_field(gpu, oracle_tp):
    g = golden("fields_gridd")
    print("[gridd] field", flush=True)
    gpu.set_field_grid_double(g["egrid"], g["bgrid"], g["lo"], g["hi"])
    bitwise(gpu.eval_field(g["pts"]), g["out"], "double grid vs reference")
    pts = _pts(2000, g["lo"], g["hi"], 6)
    f = oracle_tp.Field.grid_double(g["egrid"], g["bgrid"], g["lo"], g["hi"])
    bitwise(gpu.eval_field(pts), oracle_tp.eval_field(f, pts), "double grid vs oracle")


def _params(g, estep, pusher, precision=abi.TP_FP64):
    spec = json.loads(str(g["spec"]))
    return abi.make_params(dt=spec["dt"], estep=estep,
                           lo=(spec["xmin"], spec["ymin"], spec["zmin"]),
                           hi=(spec["xmax"], spec["ymax"], spec["zmax"]), pusher=pusher,
                           precision=precision, bc_flag=spec["bc"], nbins=spec["nbins"],
                           nt_out=spec["nt_out"], emin=spec["emin"], emax=spec["emax"])


def _compare(r, ref, what):
    for k in ("ptl", "nsteps", "espect", "espect_escape", "remain"):
        bitwise(r[k], ref[k], f"{what} {k}")
    d, dr = r["dcoeff"], ref["dcoeff"]
    assert np.all(np.abs(d - dr) <= 1e-12 * np.abs(dr) + 1e-15), what


@pytest.mark.parametrize("pusher", [abi.TP_PUSH_WIRZ, abi.TP_PUSH_BORIS])
@pytest.mark.parametrize("cells", [abi.TP_CELLS_OFF, abi.TP_CELLS_ON])
def test_mhd_track(gpu, oracle_tp, pusher, cells):
    g = golden("track_mhd")
    ptl = init_particles(700, g["lo"], g["hi"], 0.3, seed=41)    # ragged lanes
    prm = _params(g, 600, pusher)
    gpu.set_grid_cells(cells)
    try:
        gpu.set_field_mhd(g["vgrid"], g["bgrid"], g["lo"], g["hi"])
    finally:
        gpu.set_grid_cells(abi.TP_CELLS_AUTO)
    print(f"[mhd] track pusher {pusher} cells {cells}", flush=True)
    r = gpu.track_fixed(prm, ptl)
    ref = oracle_tp.track_fixed(oracle_tp.Field.mhd(g["vgrid"], g["bgrid"], g["lo"], g["hi"]),
                                prm, ptl, nthreads=1)
    _compare(r, ref, f"mhd pusher {pusher}")


def test_mhd_track_reference_fixture(gpu):
    """The reference binary's own loop on the MHD grid (Wirz, 32 particles,
    300 steps): no sin/cos of different origin on this path except the Wirz
    rotation's; 1e-10 relative."""
    g = golden("track_mhd")
    gpu.set_field_mhd(g["vgrid"], g["bgrid"], g["lo"], g["hi"])
    spec = json.loads(str(g["spec"]))
    r = gpu.track_fixed(_params(g, int(spec["tmax"] / spec["dt"]), abi.TP_PUSH_WIRZ),
                        g["ptl_in"])
    scale = np.maximum(np.abs(g["ptl_out"]), 1.0)
    assert np.max(np.abs(r["ptl"] - g["ptl_out"]) / scale) < 1e-10
    bitwise(r["espect"], g["espect"], "mhd spectrum vs reference")
    bitwise(r["nsteps"], g["nsteps"], "mhd nsteps vs reference")


def test_mhd_fp32_form(gpu):
    g = golden("track_mhd")
    ptl = init_particles(2000, g["lo"], g["hi"], 0.3, seed=42)
    gpu.set_field_mhd(g["vgrid"], g["bgrid"], g["lo"], g["hi"])
    pts = _pts(1000, g["lo"], g["hi"], 7)
    f64 = gpu.eval_field(pts)
    f32 = gpu.eval_field(pts, abi.TP_FP32_MIXED)
    assert np.max(np.abs(f32 - f64)) < 2e-6 * np.max(np.abs(f64))
    r64 = gpu.track_fixed(_params(g, 200, abi.TP_PUSH_BORIS), ptl)
    r32 = gpu.track_fixed(_params(g, 200, abi.TP_PUSH_BORIS, abi.TP_FP32_MIXED), ptl)
    bitwise(r32["nsteps"], r64["nsteps"], "fp32 nsteps")
    pos = lambda r: r["ptl"][:, :3] + r["ptl"][:, 8:11]  # noqa: E731
    assert np.max(np.abs(pos(r32) - pos(r64))) < 1e-3


def test_double_grid_track(gpu, oracle_tp):
    g = golden("fields_gridd")
    t = golden("track_mhd")
    ptl = init_particles(500, g["lo"], g["hi"], 0.3, seed=43)
    prm = _params(t, 300, abi.TP_PUSH_BORIS)
    gpu.set_field_grid_double(g["egrid"], g["bgrid"] + np.array([0.0, 0.0, 2.0]),
                              g["lo"], g["hi"])
    print("[gridd] track", flush=True)
    r = gpu.track_fixed(prm, ptl)
    f = oracle_tp.Field.grid_double(g["egrid"], g["bgrid"] + np.array([0.0, 0.0, 2.0]),
                                    g["lo"], g["hi"])
    ref = oracle_tp.track_fixed(f, prm, ptl, nthreads=1)
    _compare(r, ref, "double grid")


def test_mhd_poincare_is_the_magnetic_grid(gpu):
    """run_type 2 with system 3 traces B only (get_bfield_mhd_test_particle,
    src/emfields.c:131-143): the map equals that of the PIC grid with the
    same B."""
    g = golden("fields_mhd")
    lo, hi = g["lo"], g["hi"]
    prm = abi.make_pmap_params(face=2, ncross=3, step=0.05, length_max=200.0, lo=lo, hi=hi)
    starts = _pts(64, lo, hi, 8)[:, :3]
    starts = np.clip(starts, lo, hi)
    gpu.set_field_mhd(g["vgrid"], g["bgrid"], lo, hi)
    a = gpu.poincare_map(prm, starts)
    gpu.set_field_grid(np.zeros_like(g["vgrid"]), g["bgrid"], lo, hi)
    b = gpu.poincare_map(prm, starts)
    for k in ("xpos", "ypos", "zpos", "ncrossed", "nsteps"):
        bitwise(a[k], b[k], f"pmap {k}")
