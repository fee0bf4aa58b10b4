"""Pin the CPU oracle (oracle/tpref.c) to golden vectors written by the
reference tracer itself (tests/golden/make_golden.py drives the reference's
own objects through oracle/ref_harness.c).

With libm sin/cos the oracle performs the reference's operations in the
reference's order, so most comparisons are bitwise."""
import json

import numpy as np
import pytest

from conftest import golden
from test_particle_amd import abi


def _bitwise(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    if not same.all():
        idx = np.argwhere(~same)[:5]
        raise AssertionError(f"{(~same).sum()} differ, e.g. {idx.tolist()}: "
                             f"{a[tuple(idx[0])]!r} vs {b[tuple(idx[0])]!r}")


def test_abc_reference_coefficients(oracle_libm):
    g = golden("fields_abc_ref")
    f = oracle_libm.Field.abc()          # A=1, B=sqrt(2/3), C=sqrt(1/3)
    _bitwise(oracle_libm.eval_field(f, g["pts"]), g["out"])


def test_abc_unit_coefficients(oracle_libm):
    g = golden("fields_abc1")
    f = oracle_libm.Field.abc(1.0, 1.0, 1.0, lam=float(g["lam"]), B0=float(g["b0"]))
    _bitwise(oracle_libm.eval_field(f, g["pts"]), g["out"])


def test_abc_velocity_grid_efield(oracle_libm):
    """E = B x v with v from the double velocity grid (x/z weight exchange,
    out-of-box wrap cells)."""
    g = golden("fields_abc_vgrid")
    f = oracle_libm.Field.abc().with_velocity_grid(g["vgrid"], g["lo"], g["hi"])
    out = oracle_libm.eval_field(f, g["pts"])
    _bitwise(out, g["out"])
    assert np.abs(out[:, 3:]).max() > 1e-3      # E is really exercised


def test_pic_grid_ref_compat(oracle_libm):
    g = golden("fields_grid")
    f = oracle_libm.Field.grid(g["egrid"], g["bgrid"], g["lo"], g["hi"],
                               abi.TP_GATHER_REF_COMPAT)
    _bitwise(oracle_libm.eval_field(f, g["pts"]), g["out"])


def test_pic_grid_textbook_differs_from_reference(oracle_libm):
    """The textbook mode must not silently equal the reference's swapped
    weights (it would if the exchange were lost)."""
    g = golden("fields_grid")
    f = oracle_libm.Field.grid(g["egrid"], g["bgrid"], g["lo"], g["hi"],
                               abi.TP_GATHER_TEXTBOOK)
    out = oracle_libm.eval_field(f, g["pts"])
    assert np.abs(out - g["out"]).max() > 1e-2


@pytest.mark.parametrize("ib", [0, 1])
def test_wlcs_fields(oracle_libm, ib):
    """Elliptic integrals by AGM here vs Fukushima's series in the
    reference: agreement to 1e-12 relative to the field scale."""
    g = golden(f"fields_wlcs_b{ib}")
    f = oracle_libm.Field.wlcs(g["rows"], ibessel=int(g["ibessel"]),
                               omega=float(g["omega"]), norm_i=float(g["norm_i"]))
    out = oracle_libm.eval_field(f, g["pts"])
    ref = g["out"]
    scale = np.abs(ref).max(axis=1, keepdims=True)
    ok = np.isfinite(ref)
    assert (np.isfinite(out) == ok).all()
    err = np.abs(out - ref)[ok] / np.broadcast_to(scale, ref.shape)[ok]
    assert err.max() < 1e-12


def test_wirz_uniform_field(oracle_libm):
    g = golden("wirz_test")
    f = oracle_libm.Field.test()
    for k in g["ks"]:
        out = oracle_libm.steps(f, g["ptl"], int(k), float(g["dt"]),
                                params=_bc_params(float(g["dt"])))
        _bitwise(out, g[f"out_{k}"])


def _bc_params(dt, lo=(0, 0, 0), hi=(2 * np.pi,) * 3):
    return abi.make_params(dt=dt, estep=1, lo=lo, hi=hi)


def test_wirz_abc_chaotic_bitwise(oracle_libm):
    """1000 Wirz steps in the chaotic ABC field, bit for bit."""
    g = golden("wirz_abc1")
    f = oracle_libm.Field.abc(1.0, 1.0, 1.0)
    for k in g["ks"]:
        out = oracle_libm.steps(f, g["ptl"], int(k), float(g["dt"]),
                                params=_bc_params(float(g["dt"])))
        _bitwise(out, g[f"out_{k}"])


def test_wirz_pic_grid(oracle_libm):
    g = golden("wirz_grid")
    f = oracle_libm.Field.grid(g["egrid"], g["bgrid"], g["lo"], g["hi"])
    prm = abi.make_params(dt=float(g["dt"]), estep=1, lo=g["lo"], hi=g["hi"])
    for k in g["ks"]:
        out = oracle_libm.steps(f, g["ptl"], int(k), float(g["dt"]), params=prm)
        _bitwise(out, g[f"out_{k}"])


def test_diff_coeff(oracle_libm):
    g = golden("diff_abc1")
    f = oracle_libm.Field.abc(1.0, 1.0, 1.0)
    _bitwise(oracle_libm.diff_coeff(f, g["p0"], g["p1"], float(g["dt"])), g["out"])


def _track(oracle, name, field):
    g = golden(name)
    spec = json.loads(str(g["spec"]))
    dt = spec["dt"]
    estep = int(spec["tmax"] / dt)
    prm = abi.make_params(dt=dt, estep=estep,
                          lo=(spec["xmin"], spec["ymin"], spec["zmin"]),
                          hi=(spec["xmax"], spec["ymax"], spec["zmax"]),
                          bc_flag=spec["bc"], nbins=spec["nbins"],
                          nt_out=spec["nt_out"], emin=spec["emin"],
                          emax=spec["emax"])
    r = oracle.track_fixed(field(g), prm, g["ptl_in"], nthreads=1)
    _bitwise(r["ptl"], g["ptl_out"])
    _bitwise(r["nsteps"], g["nsteps"])
    _bitwise(r["espect"], g["espect"])
    _bitwise(r["espect_escape"], g["espect_escape"])
    _bitwise(r["dcoeff"], g["dcoeff"])
    _bitwise(r["remain"], g["remain"])
    return g, r


def test_track_loop_abc(oracle_libm):
    g, r = _track(oracle_libm, "track_abc1", lambda g: oracle_libm.Field.abc(1.0, 1.0, 1.0))
    assert r["espect"].sum() == 10 * 256
    assert (r["dcoeff"][:, 1:] != 0).all() and (r["dcoeff"][:, 0] == 0).all()


def test_track_loop_open_boundary(oracle_libm):
    """Open boundaries: escape spectrum, remaining counts, and the stale D
    terms the reference adds for escaped particles (reproduced here with
    the oracle's single-thread carry-over switch)."""
    oracle_libm.set_carry_stale(True)
    try:
        g, r = _track(oracle_libm, "track_open", lambda g: oracle_libm.Field.test())
    finally:
        oracle_libm.set_carry_stale(False)
    assert r["espect_escape"].sum() > 0
    assert r["remain"].min() < 256


def test_track_loop_pic_grid(oracle_libm):
    _track(oracle_libm, "track_grid",
           lambda g: oracle_libm.Field.grid(g["egrid"], g["bgrid"], g["lo"], g["hi"]))


def traj_layout(n, estep, nsteps_output, ptl):
    """Slots as the golden harness lays them out: estep/nsteps_output + 1
    per particle, slot 0 the initial state."""
    slots = estep // nsteps_output + 1
    accum = (np.arange(1, n + 1) * slots).astype(np.int32)
    ptl_time = np.zeros((n * slots, 11))
    ptl_time[::slots] = ptl
    return slots, accum, ptl_time


def _traj(oracle, name, field):
    g = golden(name)
    spec = json.loads(str(g["spec"]))
    dt = spec["dt"]
    estep = int(spec["tmax"] / dt)
    prm = abi.make_params(dt=dt, estep=estep,
                          lo=(spec["xmin"], spec["ymin"], spec["zmin"]),
                          hi=(spec["xmax"], spec["ymax"], spec["zmax"]),
                          bc_flag=spec["bc"], nbins=spec["nbins"],
                          nt_out=spec["nt_out"], emin=spec["emin"],
                          emax=spec["emax"])
    ptl = g["ptl_in"]
    n = ptl.shape[0]
    nout = int(spec["nsteps_output"])
    slots, accum, ptl_time = traj_layout(n, estep, nout, ptl)
    r = oracle.track_fixed_traj(field(g), prm, ptl, nout, accum, ptl_time, nthreads=1)
    _bitwise(r["ptl"], g["ptl_out"])
    _bitwise(r["ptl_time"].reshape(n, slots, 11), g["traj"])
    _bitwise(r["points"], g["points"])
    _bitwise(r["espect"], g["espect"])
    return g, r


def test_trajectory_diagnostics_pic_grid(oracle_libm):
    """traj_diagnose = 1 (src/tracking.c:440-447): snapshots every 7 steps
    of a 300-step PIC-grid run, the reference's records bit for bit."""
    g, r = _traj(oracle_libm, "traj_grid",
                 lambda g: oracle_libm.Field.grid(g["egrid"], g["bgrid"], g["lo"], g["hi"]))
    assert np.all(r["points"] == 300 // 7 + 1)


def test_trajectory_diagnostics_open_boundary(oracle_libm):
    """Open boundaries: escaped particles stop recording."""
    g, r = _traj(oracle_libm, "traj_open", lambda g: oracle_libm.Field.test())
    assert r["points"].min() < r["points"].max()


def test_pic_grid_two_time_slices(oracle_libm):
    """multi_tframe = 1: each slice summed in double, blended by the
    fraction of t / grid_dt, returned as float -- the reference's values bit
    for bit at points and times spread over several slice intervals."""
    g = golden("fields_grid2")
    f = oracle_libm.Field.grid2(g["ea"], g["ba"], g["eb"], g["bb"], g["lo"], g["hi"],
                                float(g["grid_dt"]))
    _bitwise(oracle_libm.eval_field(f, g["pts"]), g["out"])
    # at t = 0 and t = grid_dt (dt = 0) the next slice has no weight
    pts = g["pts"][:20].copy()
    z = np.zeros_like(g["ea"])
    f0 = oracle_libm.Field.grid2(z, z, g["eb"], g["bb"], g["lo"], g["hi"],
                                 float(g["grid_dt"]))
    for t in (0.0, float(g["grid_dt"])):
        pts[:, 3] = t
        _bitwise(oracle_libm.eval_field(f, pts), oracle_libm.eval_field(f0, pts))


def test_track_loop_pic_grid_two_time_slices(oracle_libm):
    _track(oracle_libm, "track_grid2",
           lambda g: oracle_libm.Field.grid2(g["ea"], g["ba"], g["eb"], g["bb"], g["lo"],
                                             g["hi"], float(g["grid_dt"])))


# --- system 3 (MHD velocity and magnetic grids) and double PIC grids ---------
def _mhd(oracle, g):
    return oracle.Field.mhd(g["vgrid"], g["bgrid"], g["lo"], g["hi"])


def test_mhd_fields(oracle_libm):
    """getemf_mhd_test_particle (src/emfields.c:107-120): float velocity and
    magnetic gathers, E = B x v in double, bitwise the reference."""
    g = golden("fields_mhd")
    out = oracle_libm.eval_field(_mhd(oracle_libm, g), g["pts"])
    _bitwise(out, g["out"])
    assert np.abs(out[:, 3:]).max() > 1e-2      # E is really exercised
    # E is perpendicular to B (E = B x v): B . E vanishes to rounding
    bdote = np.sum(out[:, :3] * out[:, 3:], axis=1)
    assert np.max(np.abs(bdote)) < 1e-12 * np.max(np.abs(out)) ** 2 * 10


def test_mhd_wirz(oracle_libm):
    g = golden("wirz_mhd")
    f = _mhd(oracle_libm, g)
    prm = abi.make_params(dt=float(g["dt"]), estep=1, lo=g["lo"], hi=g["hi"])
    for k in g["ks"]:
        out = oracle_libm.steps(f, g["ptl"], int(k), float(g["dt"]), params=prm)
        _bitwise(out, g[f"out_{k}"])


def test_track_loop_mhd(oracle_libm):
    _track(oracle_libm, "track_mhd", lambda g: _mhd(oracle_libm, g))


def test_double_grid_gathers(oracle_libm):
    """get_double_efield_at_point / get_double_bfield_at_point
    (src/electric_field.c:370-406, src/magnetic_field.c:555-591), called
    directly by the harness: bitwise."""
    g = golden("fields_gridd")
    f = oracle_libm.Field.grid_double(g["egrid"], g["bgrid"], g["lo"], g["hi"])
    _bitwise(oracle_libm.eval_field(f, g["pts"]), g["out"])


def test_double_grid_is_exact_on_linear_fields_inside_cells(oracle_libm):
    """A closed form independent of the reference: with the x and z
    fractions exchanged, the double gather of f = a + b.x returns
    a + b_x z_frac + b_y y_frac + b_z x_frac pieces; on a field that varies
    only along y (b_x = b_z = 0) it is exact."""
    nz, ny, nx = 4, 5, 6
    lo, hi = (0.0, 0.0, 0.0), (5.0, 4.0, 3.0)
    y = np.linspace(lo[1], hi[1], ny)
    e = np.zeros((nz, ny, nx, 3))
    e[..., 0] = (2.0 + 0.5 * y)[None, :, None]
    b = np.ones((nz, ny, nx, 3))
    f = oracle_libm.Field.grid_double(e, b, lo, hi)
    rng = np.random.default_rng(3)
    pts = np.column_stack([rng.uniform(0.1, 4.9, 50), rng.uniform(0.1, 3.9, 50),
                           rng.uniform(0.1, 2.9, 50), np.zeros(50)])
    out = oracle_libm.eval_field(f, pts)
    assert np.allclose(out[:, 3], 2.0 + 0.5 * pts[:, 1], rtol=0, atol=1e-14)
    assert np.allclose(out[:, :3], 1.0, rtol=0, atol=1e-14)
