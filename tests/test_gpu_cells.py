"""The cell-major copy of PIC grids (tp_gpu_set_grid_cells).

With the copy, a particle entering a cell reads that cell's 192-byte record
(the reference gather's eight corners) instead of eight nodes; the gathered
floats are the same, so every result must be bit-identical with and without
it: FP64 and FP32 forms, both pushers, periodic, open and no boundaries
(particles outside the box use the wrap cell, and points on the upper face
the clamped node path), one and two time slices -- and equal to the oracle
(the FP64 form)."""
import numpy as np
import pytest

from conftest import golden
from test_particle_amd import abi
from test_particle_amd.particles import init_particles

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def gpu():
    from test_particle_amd.gpu import Tracker
    t = Tracker(0)
    yield t
    t.set_grid_cells(abi.TP_CELLS_AUTO)
    t.close()


def same(a, b, what):
    for k in ("ptl", "nsteps", "espect", "espect_escape", "remain", "dcoeff"):
        assert np.array_equal(a[k], b[k]), f"{what}: {k} differs"


def run_both(gpu, set_field, prm, ptl):
    out = []
    for mode in (abi.TP_CELLS_OFF, abi.TP_CELLS_ON):
        gpu.set_grid_cells(mode)
        set_field()
        out.append(gpu.track_fixed(prm, ptl))
    gpu.set_grid_cells(abi.TP_CELLS_AUTO)
    return out


# the FP32 form is Boris only
@pytest.mark.parametrize("precision,pusher", [(abi.TP_FP64, abi.TP_PUSH_BORIS),
                                              (abi.TP_FP64, abi.TP_PUSH_WIRZ),
                                              (abi.TP_FP32_MIXED, abi.TP_PUSH_BORIS)])
@pytest.mark.parametrize("bc", [abi.TP_BC_PERIODIC, abi.TP_BC_OPEN, abi.TP_BC_NONE])
def test_cells_same_results(gpu, precision, pusher, bc):
    case(f"cells prec={precision} pusher={pusher} bc={bc}")
    g = golden("track_grid")
    ptl = init_particles(700, g["lo"], g["hi"], 0.5, seed=71)
    # a few particles on the upper faces, on nodes and outside the box
    lo, hi = np.asarray(g["lo"], float), np.asarray(g["hi"], float)
    ptl[:3, 0:3] = hi
    ptl[3, 0:3] = lo
    ptl[4, 0:3] = hi + 0.3
    ptl[5, 0:3] = lo - 0.3
    prm = abi.make_params(dt=0.02, estep=400, lo=lo, hi=hi, pusher=pusher, precision=precision,
                          bc_flag=bc, nbins=50, nt_out=4, emin=1e-3, emax=1.0)
    off, on = run_both(gpu, lambda: gpu.set_field_grid(g["egrid"], g["bgrid"], lo, hi), prm,
                       ptl)
    same(off, on, "cells vs nodes")
    if bc == abi.TP_BC_NONE:
        # particles left the box: the wrap and clamped cells were used
        p = off["ptl"]
        assert ((p[:, 0:3] > hi) | (p[:, 0:3] < lo)).any()


def test_cells_equal_oracle(gpu, oracle_tp):
    case("cells vs oracle")
    g = golden("track_grid")
    ptl = init_particles(300, g["lo"], g["hi"], 0.3, seed=72)
    lo, hi = np.asarray(g["lo"], float), np.asarray(g["hi"], float)
    prm = abi.make_params(dt=0.01, estep=300, lo=lo, hi=hi, pusher=abi.TP_PUSH_BORIS,
                          nbins=50, nt_out=3, emin=1e-3, emax=1.0)
    gpu.set_grid_cells(abi.TP_CELLS_ON)
    gpu.set_field_grid(g["egrid"], g["bgrid"], lo, hi)
    r = gpu.track_fixed(prm, ptl)
    gpu.set_grid_cells(abi.TP_CELLS_AUTO)
    f = oracle_tp.Field.grid(g["egrid"], g["bgrid"], lo, hi, abi.TP_GATHER_REF_COMPAT)
    ref = oracle_tp.track_fixed(f, prm, ptl, nthreads=1)
    for k in ("ptl", "nsteps", "espect", "espect_escape", "remain"):
        assert np.array_equal(r[k], ref[k]), k


@pytest.mark.parametrize("precision", [abi.TP_FP64, abi.TP_FP32_MIXED])
def test_cells_two_slices(gpu, precision):
    case(f"cells two slices prec={precision}")
    g = golden("track_grid2")
    lo, hi = np.asarray(g["lo"], float), np.asarray(g["hi"], float)
    prm = abi.make_params(dt=float(g["dt"]), estep=300, lo=lo, hi=hi,
                          pusher=abi.TP_PUSH_BORIS, precision=precision, nbins=50, nt_out=3,
                          emin=1e-3, emax=1.0)

    def setf():
        gpu.set_field_grid2(g["ea"], g["ba"], g["eb"], g["bb"], lo, hi, float(g["grid_dt"]))

    off, on = run_both(gpu, setf, prm, g["ptl_in"])
    same(off, on, "two-slice cells vs nodes")


def test_cells_mode_errors(gpu):
    from test_particle_amd.gpu import TPError
    for bad in (-1, 3):
        with pytest.raises(TPError):
            gpu.set_grid_cells(bad)


def _mu(b, u):
    return (b * u).sum(1) / (np.sqrt((b * b).sum(1)) * np.sqrt((u * u).sum(1)))


@pytest.mark.parametrize("two", [False, True])
@pytest.mark.parametrize("cells", [abi.TP_CELLS_OFF, abi.TP_CELLS_ON])
def test_fp32_form_d_terms_use_the_fp64_field(gpu, two, cells):
    """The FP32 form's D terms take the field at the particle from the FP64
    reference gather (served from the FP32 cell cache when that holds the
    same cell): D_mumu of the last row equals the sum of (mu - mu0)^2 / t
    over the returned particles with mu from tp_gpu_eval_field (FP64) at
    their final state, to summation order."""
    case(f"FP32 D terms two={two} cells={cells}")
    g = golden("track_grid2" if two else "track_grid")
    lo, hi = np.asarray(g["lo"], float), np.asarray(g["hi"], float)
    ptl = init_particles(1000, lo, hi, 0.5, seed=73)
    gpu.set_grid_cells(cells)
    if two:
        gpu.set_field_grid2(g["ea"], g["ba"], g["eb"], g["bb"], lo, hi, float(g["grid_dt"]))
    else:
        gpu.set_field_grid(g["egrid"], g["bgrid"], lo, hi)
    gpu.set_grid_cells(abi.TP_CELLS_AUTO)
    dt = 0.02
    prm = abi.make_params(dt=dt, estep=200, lo=lo, hi=hi, pusher=abi.TP_PUSH_BORIS,
                          precision=abi.TP_FP32_MIXED, nbins=50, nt_out=2, emin=1e-3,
                          emax=1.0)
    r = gpu.track_fixed(prm, ptl)
    p = r["ptl"]
    assert r["remain"][2] == len(ptl)
    t = p[:, 7] * 2 + dt
    f0 = gpu.eval_field(np.column_stack([ptl[:, 0:3], ptl[:, 7]]))
    f1 = gpu.eval_field(np.column_stack([p[:, 0:3], p[:, 7]]))
    mu0 = _mu(f0[:, 0:3], ptl[:, 4:7])
    mu1 = _mu(f1[:, 0:3], p[:, 4:7])
    duu = ((mu1 - mu0) ** 2 / t).sum()
    assert abs(r["dcoeff"][5, 2] / duu - 1) < 1e-12
    dxx = ((ptl[:, 0] - (p[:, 0] + p[:, 8])) ** 2 / t).sum()
    assert abs(r["dcoeff"][1, 2] / dxx - 1) < 1e-12
