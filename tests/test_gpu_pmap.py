"""Poincare maps on the GPU (tp_gpu_poincare_map, csrc/tp_fieldline.hip).

* the reference's own crossing points (golden vectors of
  src/poincare_map.c's RK4 step and config parser on PIC grids) bit for
  bit, through any chunking of the steps into launches, with one lane per
  line and with four (one per component, the form of the launches over few
  running lines);
* the CPU oracle bit for bit for the analytic fields (ABC, multi-mode,
  with the shared sin/cos) and the textbook PIC gather, in both crossing
  modes, on lines that end at different times (the active-list compaction
  between launches);
* a 65,536-line map of a turbulent PIC grid: every periodic-mode record lies
  within one step of its plane, the run is deterministic, and a sample of
  lines equals the oracle bit for bit;
* argument and state errors.
"""
import json

import numpy as np
import pytest

from conftest import golden
from test_particle_amd import abi
from test_particle_amd.gpu import pmap_starts

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * np.pi


def case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def gpu():
    from test_particle_amd.gpu import Tracker
    t = Tracker(0)
    yield t
    t.close()


def bitwise(a, b, what=""):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    if bad.any():
        i = tuple(np.argwhere(bad)[0])
        raise AssertionError(f"{what}: {bad.sum()} differ, first {i}: {a[i]!r} vs {b[i]!r}")


def same_map(r, ref, what):
    for k in ("xpos", "ypos", "zpos"):
        bitwise(r[k], ref[k], f"{what} {k}")
    assert np.array_equal(r["ncrossed"], ref["ncrossed"]), what
    if "nsteps" in ref:
        assert np.array_equal(r["nsteps"], ref["nsteps"]), what


def _golden_case(gpu, name):
    g = golden(name)
    c = json.loads(str(g["cfg"]))
    if name == "pmap_grid2":
        z = np.zeros_like(g["b_next"])
        gpu.set_field_grid2(z, g["b_next"], z, g["b_prev"], g["lo"], g["hi"],
                            float(g["grid_dt"]))
    else:
        gpu.set_field_grid(np.zeros_like(g["bgrid"]), g["bgrid"], g["lo"], g["hi"])
    st = pmap_starts(c["face"], (c["x0"], c["y0"], c["z0"]), single=bool(c["single"]),
                     min_h=c["min_h"], max_h=c["max_h"], min_v=c["min_v"], max_v=c["max_v"],
                     nh=c["nh"], nv=c["nv"])
    return g, c, st


@pytest.mark.parametrize("name", ["pmap_grid", "pmap_single", "pmap_grid2"])
@pytest.mark.parametrize("chunk,lanes", [(0, 0), (37, 0), (0, 1), (37, 1), (37, 4)])
def test_reference_golden_maps(gpu, name, chunk, lanes):
    case(f"poincare golden {name}, chunk {chunk}, lanes {lanes}")
    g, c, st = _golden_case(gpu, name)
    prm = abi.make_pmap_params(face=c["face"], ncross=c["ncross"], step=c["step"],
                               length_max=c["length_max"], lo=g["lo"], hi=g["hi"], chunk=chunk,
                               lanes=lanes)
    r = gpu.poincare_map(prm, st)
    ref = dict(xpos=g["xpos"], ypos=g["ypos"], zpos=g["zpos"],
               ncrossed=(g["xpos"] != 2 * g["hi"][0]).sum(axis=1) - 1)
    same_map(r, ref, name)
    if chunk:
        assert gpu.last_kernel_time()[1] > 1


def _starts(n, lo, hi, face, seed):
    rng = np.random.default_rng(seed)
    st = np.column_stack([rng.uniform(lo[d], hi[d], n) for d in range(3)])
    st[:, face] = rng.uniform(lo[face], hi[face])          # one plane
    return st


FIELDS = ["abc", "modes", "grid_tb", "grid_ref"]


def _set(gpu, oracle, kind, seed=0):
    if kind == "abc":
        gpu.set_field_abc(1.0, 1.0, 1.0)
        return oracle.Field.abc(1.0, 1.0, 1.0), (0.0,) * 3, (TWO_PI,) * 3
    if kind == "modes":
        from test_particle_amd.fields import mode_table
        rows = mode_table(48, kmin=0.5, kmax=8.0, rms=0.8, seed=5)
        gpu.set_field_modes(rows, 1.0)
        L = TWO_PI / 0.5
        return oracle.Field.modes(rows, 1.0), (0.0,) * 3, (L,) * 3
    rng = np.random.default_rng(40 + seed)
    lo, hi = (0.0, -1.0, 2.0), (6.0, 4.0, 5.0)
    b = (0.6 * rng.normal(0, 1, (8, 10, 12, 3))).astype(np.float32)
    b[..., 2] += 0.5
    e = np.zeros_like(b)
    mode = abi.TP_GATHER_TEXTBOOK if kind == "grid_tb" else abi.TP_GATHER_REF_COMPAT
    gpu.set_field_grid(e, b, lo, hi, mode)
    return oracle.Field.grid(e, b, lo, hi, mode), lo, hi


@pytest.mark.parametrize("kind", FIELDS + ["grid_ref_1lane"])
@pytest.mark.parametrize("mode", [abi.TP_PMAP_REF_COMPAT, abi.TP_PMAP_PERIODIC])
def test_matches_oracle(gpu, oracle_tp, kind, mode):
    lanes = 1 if kind.endswith("_1lane") else 0
    kind = kind.replace("_1lane", "")
    f, lo, hi = _set(gpu, oracle_tp, kind)
    face = 2 if kind.startswith("grid") else 1
    n = 300
    ncross, lmax = {"abc": (6, 25.0), "modes": (6, 60.0)}.get(kind, (16, 60.0))
    st = _starts(n, lo, hi, face, seed=7)
    step = 0.03
    case(f"poincare {kind} mode {mode}: {n} lines, ncross {ncross}, length {lmax}")
    prm = abi.make_pmap_params(face=face, ncross=ncross, step=step, length_max=lmax, lo=lo,
                               hi=hi, mode=mode, chunk=200, lanes=lanes)
    r = gpu.poincare_map(prm, st)
    ref = oracle_tp.poincare_map(f, prm, st, nthreads=8)
    same_map(r, ref, f"{kind} mode {mode}")
    # lines end at different times: some reach ncross, some the length cap
    nc = r["ncrossed"]
    assert (nc == ncross).any() and (nc < ncross).any(), np.bincount(nc)


def _set_more(gpu, oracle, kind, ibessel=1):
    rng = np.random.default_rng(61)
    lo, hi = (0.0, -1.0, 2.0), (6.0, 4.0, 5.0)
    if kind == "test":
        gpu.set_field_test()
        return oracle.Field.test(), lo, hi
    if kind == "wlcs":
        g = golden(f"fields_wlcs_b{ibessel}")
        kw = dict(ibessel=int(g["ibessel"]), omega=float(g["omega"]), norm_i=float(g["norm_i"]))
        gpu.set_field_wlcs(g["rows"], **kw)
        return oracle.Field.wlcs(g["rows"], **kw), (-1.0,) * 3, (2.0,) * 3
    b = [(0.6 * rng.normal(0, 1, (8, 10, 12, 3))).astype(np.float32) for _ in range(2)]
    b[0][..., 0] += 0.5                      # next slice: unused at t = 0
    b[1][..., 2] += 0.5                      # previous slice
    e = np.zeros_like(b[0])
    mode = abi.TP_GATHER_TEXTBOOK if kind == "grid2_tb" else abi.TP_GATHER_REF_COMPAT
    gpu.set_field_grid2(e, b[0], e, b[1], lo, hi, 0.7, mode)
    return oracle.Field.grid2(e, b[0], e, b[1], lo, hi, 0.7, mode), lo, hi


@pytest.mark.parametrize("kind", ["grid2_ref", "grid2_tb", "test"])
@pytest.mark.parametrize("mode", [abi.TP_PMAP_REF_COMPAT, abi.TP_PMAP_PERIODIC])
def test_matches_oracle_more_fields(gpu, oracle_tp, kind, mode):
    """Two time slices at t = 0 (only the previous slice enters) and the
    uniform test field (straight lines along z), bit for bit."""
    f, lo, hi = _set_more(gpu, oracle_tp, kind)
    st = _starts(300, lo, hi, 2, seed=13)
    prm = abi.make_pmap_params(face=2, ncross=12, step=0.03, length_max=40.0, lo=lo, hi=hi,
                               mode=mode, chunk=300)
    case(f"poincare {kind} mode {mode}: 300 lines")
    r = gpu.poincare_map(prm, st)
    ref = oracle_tp.poincare_map(f, prm, st, nthreads=8)
    same_map(r, ref, f"{kind} mode {mode}")
    if kind == "test":
        # straight lines along z: x and y never change, one crossing per
        # period (periodic test) or per fold and per period (reference's)
        assert (r["xpos"][:, 1:][r["xpos"][:, 1:] != 2 * hi[0]] == np.repeat(
            st[:, :1], 12, axis=1)[r["xpos"][:, 1:] != 2 * hi[0]]).all()
        per = int(np.floor(40.0 / (hi[2] - lo[2])))
        expect = min(12, per if mode == abi.TP_PMAP_PERIODIC else 2 * per + 1)
        assert np.all(r["ncrossed"] == expect), np.bincount(r["ncrossed"])


@pytest.mark.parametrize("ibessel,tol", [(0, 1e-9), (1, 1e-6)])
def test_wlcs_lines_follow_the_oracle(gpu, oracle_tp, ibessel, tol):
    """The WLCS field differs from the oracle's by ~1e-12 with static wires
    and ~2e-6 with the float Bessel wires (the device math library's acos,
    j0f and j1f), so crossing points agree to a tolerance rather than bit for
    bit."""
    f, lo, hi = _set_more(gpu, oracle_tp, "wlcs", ibessel)
    st = _starts(200, lo, hi, 0, seed=17)
    prm = abi.make_pmap_params(face=0, ncross=6, step=0.01, length_max=12.0, lo=lo, hi=hi,
                               mode=abi.TP_PMAP_PERIODIC)
    case(f"poincare wlcs ibessel {ibessel}: 200 lines")
    r = gpu.poincare_map(prm, st)
    ref = oracle_tp.poincare_map(f, prm, st, nthreads=8)
    same = r["ncrossed"] == ref["ncrossed"]
    assert same.mean() > 0.98
    for k in ("xpos", "ypos", "zpos"):
        d = np.abs(r[k][same] - ref[k][same])
        assert d.max() < tol, (k, d.max())


def _fp32_vs_fp64(gpu, prm, st):
    prm.precision = abi.TP_FP64
    r64 = gpu.poincare_map(prm, st)
    prm.precision = abi.TP_FP32_MIXED
    r32 = gpu.poincare_map(prm, st)
    prm.precision = abi.TP_FP64
    return r64, r32


def test_fp32_form_straight_lines(gpu):
    """Uniform field on a grid (straight lines): the FP32 form records the
    same crossings as the FP64 form, positions within 1e-5."""
    lo, hi = (0.0, 0.0, 0.0), (4.0, 3.0, 5.0)
    b = np.broadcast_to(np.array([0.3, 0.8, 0.5], np.float32), (5, 5, 5, 3)).copy()
    gpu.set_field_grid(np.zeros_like(b), b, lo, hi)
    st = _starts(500, lo, hi, 1, seed=19)
    prm = abi.make_pmap_params(face=1, ncross=10, step=0.05, length_max=60.0, lo=lo, hi=hi,
                               mode=abi.TP_PMAP_PERIODIC)
    case("poincare FP32 form, uniform grid field")
    r64, r32 = _fp32_vs_fp64(gpu, prm, st)
    assert np.array_equal(r32["ncrossed"], r64["ncrossed"])
    assert np.array_equal(r32["nsteps"], r64["nsteps"])
    for k in ("xpos", "ypos", "zpos"):
        assert np.abs(r32[k] - r64[k]).max() < 1e-5, k


@pytest.mark.parametrize("kind", ["grid_ref", "grid_tb", "abc", "modes"])
def test_fp32_form_first_crossings(gpu, oracle_tp, kind):
    """Turbulent and chaotic fields: the FP32 form's first crossing lies
    within 1e-3 of the FP64 one for 95 % of the lines (the FP64 form is the
    oracle's bit for bit)."""
    f, lo, hi = _set(gpu, oracle_tp, kind)
    face = 2 if kind.startswith("grid") else 1
    st = _starts(2000, lo, hi, face, seed=23)
    prm = abi.make_pmap_params(face=face, ncross=3, step=0.02, length_max=60.0, lo=lo, hi=hi,
                               mode=abi.TP_PMAP_PERIODIC)
    case(f"poincare FP32 form vs FP64, {kind}")
    r64, r32 = _fp32_vs_fp64(gpu, prm, st)
    both = (r64["ncrossed"] >= 1) & (r32["ncrossed"] >= 1)
    assert both.mean() > 0.5
    d = np.sqrt(sum((r32[k][both, 1] - r64[k][both, 1]) ** 2 for k in ("xpos", "ypos", "zpos")))
    print(f"first-crossing distance FP32 vs FP64: median {np.median(d):.2e}, "
          f"p95 {np.quantile(d, 0.95):.2e}; same count {np.mean(r32['ncrossed'] == r64['ncrossed']):.3f}")
    assert np.quantile(d, 0.95) < 1e-3
    assert np.mean(r32["ncrossed"] == r64["ncrossed"]) > 0.9


def test_chunking_does_not_change_results(gpu, oracle_tp):
    _set(gpu, oracle_tp, "abc")
    st = _starts(1000, (0.0,) * 3, (TWO_PI,) * 3, 0, seed=9)
    out = []
    for chunk in (1, 13, 0):
        case(f"poincare abc chunk {chunk}")
        prm = abi.make_pmap_params(face=0, ncross=6, step=0.05, length_max=15.0,
                                   lo=(0.0,) * 3, hi=(TWO_PI,) * 3,
                                   mode=abi.TP_PMAP_PERIODIC, chunk=chunk)
        out.append(gpu.poincare_map(prm, st))
        ms, launches = gpu.last_kernel_time()
        if chunk == 1:
            assert launches == out[-1]["nsteps"].max()
    same_map(out[0], out[1], "chunk 1 vs 13")
    same_map(out[0], out[2], "chunk 1 vs default")


def test_turbulent_grid_map_at_scale(gpu, oracle_tp):
    from test_particle_amd.fields import pic_turbulence_grid
    e, b, lo, hi = pic_turbulence_grid(n=64, seed=3, db_rms=0.6)
    gpu.set_field_grid(np.zeros_like(b), b, lo, hi)
    st = pmap_starts(2, (0.0, 0.0, 17.25), min_h=0.25, max_h=63.75, min_v=0.25, max_v=63.75,
                     nh=256, nv=256)
    step, ncross = 0.1, 6
    prm = abi.make_pmap_params(face=2, ncross=ncross, step=step, length_max=600.0, lo=lo, hi=hi,
                               mode=abi.TP_PMAP_PERIODIC)
    case(f"poincare turbulent 64^3 grid: {st.shape[0]} lines, ncross {ncross}")
    r = gpu.poincare_map(prm, st)
    r2 = gpu.poincare_map(prm, st)
    same_map(r, r2, "rerun")
    nc = r["ncrossed"]
    print(f"crossings per line: {np.bincount(nc, minlength=ncross + 1).tolist()}")
    # a line that wanders far from the guide direction can miss its plane
    assert (nc >= 1).mean() > 0.99 and nc.max() == ncross
    L = hi[2] - lo[2]
    for i in range(0, st.shape[0], 997):
        z = r["zpos"][i, 1:nc[i] + 1]
        off = (z - st[i, 2] + 0.5 * L) % L - 0.5 * L
        assert np.abs(off).max() <= step, i
        assert (r["zpos"][i, nc[i] + 1:] == 2 * hi[2]).all()
    # every 1365th line plus lines of each crossing count (the rare ones
    # included)
    sample = set(range(0, st.shape[0], st.shape[0] // 48))
    for k in range(ncross + 1):
        sample.update(np.flatnonzero(nc == k)[:4].tolist())
    sample = np.array(sorted(sample))
    case(f"oracle on {sample.size} sample lines")
    ref = oracle_tp.poincare_map(oracle_tp.Field.grid(np.zeros_like(b), b, lo, hi), prm,
                                 st[sample], nthreads=8)
    sub = {k: r[k][sample] for k in ("xpos", "ypos", "zpos", "ncrossed", "nsteps")}
    same_map(sub, ref, "sample")


def test_errors(gpu):
    from test_particle_amd.gpu import TPError, Tracker
    st = np.array([[0.1, 0.2, 0.3]])
    gpu.set_field_abc()
    for kw in (dict(face=-1), dict(step=-0.1), dict(length_max=np.nan), dict(mode=5),
               dict(chunk=-2), dict(hi=(0.0, 1.0, 1.0)), dict(precision=3), dict(lanes=2)):
        args = dict(face=0, ncross=2, step=0.1, length_max=5.0, lo=(0.0,) * 3, hi=(1.0,) * 3)
        args.update(kw)
        with pytest.raises(TPError) as ei:
            gpu.poincare_map(abi.make_pmap_params(**args), st)
        assert ei.value.code == abi.TP_E_ARG, kw
    with Tracker(0) as fresh:
        with pytest.raises(TPError) as ei:
            fresh.poincare_map(abi.make_pmap_params(face=0, ncross=2, step=0.1, length_max=5.0,
                                                    lo=(0.0,) * 3, hi=(1.0,) * 3), st)
        assert ei.value.code == abi.TP_E_STATE
    # no lines: nothing to do
    r = gpu.poincare_map(abi.make_pmap_params(face=0, ncross=2, step=0.1, length_max=5.0,
                                              lo=(0.0,) * 3, hi=(1.0,) * 3), np.zeros((0, 3)))
    assert r["xpos"].shape == (0, 3)
