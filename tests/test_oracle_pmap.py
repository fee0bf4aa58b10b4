"""Poincare map (run_type 2, src/poincare_map.c) in the CPU oracle.

Pins:
* golden vectors of the reference itself: its config parser and RK4 step
  (config_poincare_map, next_location) driven through oracle/ref_harness.c
  (mode pmap) on PIC grids -- one time slice, two time slices at t = 0, a
  grid of start points and a single point on a box face; the oracle's
  crossing points must equal the reference's bit for bit, and the start
  points come from the library's host helper tp_host_pmap_starts;
* closed forms in a uniform field, where field lines are straight: with
  the periodic crossing test the k-th record is at the first RK4 step whose
  arc length reaches k L / b_face; with the reference's test the records are
  the first step, the steps that fold the face axis and the true plane
  crossings, merged in order;
* argument checks.
"""
import json
import math

import numpy as np
import pytest

from conftest import golden
from test_particle_amd import abi
from test_particle_amd.gpu import pmap_starts


def _case(name):
    g = golden(name)
    c = json.loads(str(g["cfg"]))
    prm = abi.make_pmap_params(face=c["face"], ncross=c["ncross"], step=c["step"],
                               length_max=c["length_max"], lo=g["lo"], hi=g["hi"])
    st = pmap_starts(c["face"], (c["x0"], c["y0"], c["z0"]), single=bool(c["single"]),
                     min_h=c["min_h"], max_h=c["max_h"], min_v=c["min_v"], max_v=c["max_v"],
                     nh=c["nh"], nv=c["nv"])
    return g, c, prm, st


def _field(oracle, name, g):
    if name == "pmap_grid2":
        z = np.zeros_like(g["b_next"])
        return oracle.Field.grid2(z, g["b_next"], z, g["b_prev"], g["lo"], g["hi"],
                                  float(g["grid_dt"]))
    return oracle.Field.grid(np.zeros_like(g["bgrid"]), g["bgrid"], g["lo"], g["hi"])


def _bitwise(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    assert not bad.any(), (bad.sum(), np.argwhere(bad)[:3].tolist())


@pytest.mark.parametrize("name", ["pmap_grid", "pmap_single", "pmap_grid2"])
def test_poincare_map_matches_reference(oracle_libm, name):
    g, c, prm, st = _case(name)
    # the host helper lays the start points out as the reference does
    _bitwise(st, np.stack([g["xpos"][:, 0], g["ypos"][:, 0], g["zpos"][:, 0]], axis=1))
    r = oracle_libm.poincare_map(_field(oracle_libm, name, g), prm, st)
    _bitwise(r["xpos"], g["xpos"])
    _bitwise(r["ypos"], g["ypos"])
    _bitwise(r["zpos"], g["zpos"])
    filled = (g["xpos"] != 2 * g["hi"][0]).sum(axis=1) - 1
    assert np.array_equal(r["ncrossed"], filled)


def test_golden_cases_cover_both_endings():
    """One case reaches ncross on every line, two stop at the length cap."""
    ends = {}
    for name in ("pmap_grid", "pmap_single", "pmap_grid2"):
        g, c, _, _ = _case(name)
        n = (g["xpos"] != 2 * g["hi"][0]).sum(axis=1) - 1
        ends[name] = bool((n == c["ncross"]).all())
    assert ends == {"pmap_grid": False, "pmap_single": True, "pmap_grid2": False}


def test_reference_records_fold_points(oracle_libm):
    """The reference's crossing test compares the face coordinate before a
    step's periodic fold: a fold of the face axis is recorded as a crossing
    (records alternate between the box edge and the start plane), and the
    first step from the start is a record."""
    g, c, prm, st = _case("pmap_grid")
    y = g["ypos"][0]
    lo, L = g["lo"][1], g["hi"][1] - g["lo"][1]
    assert abs(y[1] - st[0, 1]) < c["step"]                # one step from the start
    near_lo = np.abs(y[2:12] - lo) < 2 * c["step"]
    near_plane = np.abs(y[2:12] - st[0, 1]) < c["step"]
    assert near_lo[::2].all() and near_plane[1::2].all()
    assert L > 10 * c["step"]


# ---------------------------------------------------------------- closed forms
LO, HI = np.array([0.0, 0.0, 0.0]), np.array([4.0, 3.0, 5.0])
B_DIR = np.array([0.3, 0.8, 0.5], np.float32)
L_Z = HI[2] - LO[2]


def _uniform_field(oracle):
    b = np.broadcast_to(B_DIR, (5, 5, 5, 3)).astype(np.float32)
    return oracle.Field.grid(np.zeros_like(b), b, LO, HI, abi.TP_GATHER_TEXTBOOK)


def _bhat():
    b = B_DIR.astype(np.float64)
    return b / np.linalg.norm(b)


def _expected_steps(mode, y0, face, step, ncross, lmax):
    """Record steps of a straight line with b_face > 0 (tie-free inputs)."""
    bh = _bhat()
    v = step * bh[face]                      # face advance per step
    L = HI[face] - LO[face]
    kmax = math.ceil(lmax / step)            # steps before the length cap
    plane = [math.ceil(j * L / v) for j in range(1, kmax)]
    for j in range(1, 6):
        assert 1e-4 < (j * L / v) % 1.0 < 1 - 1e-4          # no near-ties
    if mode == abi.TP_PMAP_PERIODIC:
        ks = plane
    else:
        for j in range(1, 6):
            assert 1e-4 < ((HI[face] + (j - 1) * L - y0) / v) % 1.0 < 1 - 1e-4
        folds = [math.floor((HI[face] + (j - 1) * L - y0) / v) + 2 for j in range(1, kmax)]
        ks = sorted(set([1] + plane + folds))
    ks = [k for k in ks if k < kmax][:ncross]
    return ks, kmax


@pytest.mark.parametrize("mode", [abi.TP_PMAP_REF_COMPAT, abi.TP_PMAP_PERIODIC])
@pytest.mark.parametrize("ncross,lmax", [(9, 500.0), (40, 30.013)])
def test_straight_lines_closed_form(oracle_libm, mode, ncross, lmax):
    face, step = 1, 0.05
    starts = np.array([[0.7, 1.37, 2.2], [3.1, 2.01, 0.4], [1.9, 0.61, 4.6]])
    prm = abi.make_pmap_params(face=face, ncross=ncross, step=step, length_max=lmax, lo=LO,
                               hi=HI, mode=mode)
    r = oracle_libm.poincare_map(_uniform_field(oracle_libm), prm, starts)
    bh = _bhat()
    L = HI - LO
    for i, s0 in enumerate(starts):
        ks, kmax = _expected_steps(mode, s0[face], face, step, ncross, lmax)
        assert r["ncrossed"][i] == len(ks)
        assert r["nsteps"][i] == (ks[-1] if len(ks) == ncross else kmax)
        rec = np.stack([r["xpos"][i], r["ypos"][i], r["zpos"][i]], axis=1)
        assert np.array_equal(rec[0], s0)
        for j, k in enumerate(ks, start=1):
            # the record is the line after k steps, up to whole periods
            d = (rec[j] - (s0 + k * step * bh)) / L
            assert np.abs(d - np.rint(d)).max() < 1e-9, (i, j, k)
        assert (rec[len(ks) + 1:] == 2 * HI).all()


def test_periodic_mode_on_a_box_face(oracle_libm):
    """A start plane on the box face z = zmin: the periodic test records one
    crossing per period along z (the reference's test misses those where
    the line leaves through zmax)."""
    face, step = 2, 0.05
    starts = np.array([[0.7, 1.37, 0.0]])
    ks, kmax = _expected_steps(abi.TP_PMAP_PERIODIC, 0.0, face, step, 6, 400.0)
    prm = abi.make_pmap_params(face=face, ncross=6, step=step, length_max=400.0, lo=LO, hi=HI,
                               mode=abi.TP_PMAP_PERIODIC)
    r = oracle_libm.poincare_map(_uniform_field(oracle_libm), prm, starts)
    assert r["ncrossed"][0] == 6 and r["nsteps"][0] == ks[-1]
    z = r["zpos"][0, 1:]
    assert (np.abs((z - LO[2] + 0.5 * L_Z) % L_Z - 0.5 * L_Z) < step).all()
    prm.mode = abi.TP_PMAP_REF_COMPAT
    r2 = oracle_libm.poincare_map(_uniform_field(oracle_libm), prm, starts)
    assert r2["ncrossed"][0] == 1      # only the first step: no later sign change


def test_zero_crossings_and_bad_arguments(oracle_libm):
    f = _uniform_field(oracle_libm)
    st = np.array([[0.7, 1.37, 2.2]])
    prm = abi.make_pmap_params(face=1, ncross=0, step=0.05, length_max=10.0, lo=LO, hi=HI)
    r = oracle_libm.poincare_map(f, prm, st)
    assert r["ncrossed"][0] == 0 and r["nsteps"][0] == 0
    assert np.array_equal(r["ypos"][0], [1.37])
    for kw in (dict(face=3), dict(step=0.0), dict(step=np.inf), dict(length_max=np.inf),
               dict(ncross=-1),
               dict(step=1e-20, length_max=1.0), dict(mode=2)):
        args = dict(face=1, ncross=3, step=0.05, length_max=10.0, lo=LO, hi=HI)
        args.update(kw)
        with pytest.raises(ValueError):
            oracle_libm.poincare_map(f, abi.make_pmap_params(**args), st)


def test_start_layout_helper():
    st = pmap_starts(0, (2.5, -1.0, 7.0), min_h=0.0, max_h=1.0, min_v=10.0, max_v=12.0, nh=3,
                     nv=2)
    # point iv*nh + ih: y = ih/2, z = 10 + 2 iv, x = the plane
    assert st.tolist() == [[2.5, 0.0, 10.0], [2.5, 0.5, 10.0], [2.5, 1.0, 10.0],
                           [2.5, 0.0, 12.0], [2.5, 0.5, 12.0], [2.5, 1.0, 12.0]]
    one = pmap_starts(1, (1.0, 2.0, 3.0), min_h=4.0, max_h=8.0, min_v=0.0, max_v=5.0, nh=1,
                      nv=1)
    # nh = nv = 1: the spacing is the whole range, the point sits at the minima
    assert one.tolist() == [[4.0, 2.0, 0.0]]
    assert pmap_starts(2, (1.0, 2.0, 3.0), single=True).tolist() == [[1.0, 2.0, 3.0]]


INIT_DAT_BLOCK = """Run type (1: test particle simulation 2: Poincare map): 2
------------------------- Poincare map configuration -------------------------
Which face as the cross section (0: x, 1: y, 2:z): 1
Field-line tracking step: 1E-2
Maximum tracking length: 2E6
Number of crossing points: 1000
Whether to get the map only for a single point: 0
x-position of the point: 10
y-position of the point: 0
z-position of the point: 0
---------- for a 2D plane --------
min_horizontal: 10
max_horizontal: 990.0
min_vertical: -0.0
max_vertical: 1.0
nhorizontal: 101
nvertical: 1
-------------------------------------------------------------------------------
"""


def test_config_block(tmp_path, oracle_libm):
    """read_config parses the reference's init.dat block (the values of its
    PIC configuration); a block written by write_config and read back
    reproduces the reference's golden map, whose values the reference's own
    config_poincare_map parsed."""
    from test_particle_amd import poincare
    p = tmp_path / "init.dat"
    p.write_text(INIT_DAT_BLOCK)
    cfg = poincare.read_config(str(p))
    assert cfg == dict(face=1, step=1e-2, length_max=2e6, ncross=1000, single=0, x0=10.0,
                       y0=0.0, z0=0.0, min_h=10.0, max_h=990.0, min_v=-0.0, max_v=1.0, nh=101,
                       nv=1)
    st = poincare.starts(cfg)
    assert st.shape == (101, 3) and st[1, 0] == 9.8 * 1 + 10.0 and (st[:, 1] == 0.0).all()
    for name in ("pmap_grid", "pmap_single", "pmap_grid2"):
        g, c, _, _ = _case(name)
        q = tmp_path / f"{name}.dat"
        poincare.write_config(str(q), c)
        q.write_text("Some other block\nwith lines\n" + q.read_text())
        back = poincare.read_config(str(q))
        r = oracle_libm.poincare_map(_field(oracle_libm, name, g),
                                     poincare.params(back, g["lo"], g["hi"]),
                                     poincare.starts(back))
        _bitwise(r["xpos"], g["xpos"])
        _bitwise(r["zpos"], g["zpos"])
    bad = tmp_path / "bad.dat"
    bad.write_text(INIT_DAT_BLOCK.replace("(0: x, 1: y, 2:z): 1", "(0: x, 1: y, 2:z): 4"))
    with pytest.raises(ValueError):
        poincare.read_config(str(bad))
