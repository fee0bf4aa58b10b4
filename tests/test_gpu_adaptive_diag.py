"""The adaptive stepper with dense output and the fixed-step loop's
diagnostics on the GPU (tp_gpu_run_adaptive_diag), against the oracle's
tpref_bs_track_diag (pinned by closed forms and DOP853 in
test_oracle_bs_dense.py).

* bitwise: particles, accepted steps, spectra, escape spectra and remaining
  counts, and the accepted/rejected/RHS totals; D sums to 1e-12 (the GPU
  sums per-particle terms as a tree, the oracle in particle order) -- in a
  uniform field with periodic wraps, the multi-mode field with periodic and
  open boundaries, and the PIC grid through the reference gather, for
  several launch shapes (other lanes take each particle);
* against the plain FP64 fixed-step kernel on the same field and output
  grid: D_xx and the spectra agree statistically (trajectories of the two
  integrators decorrelate in the turbulent field);
* trajectory records (tp_gpu_track_adaptive_diag_traj) bitwise against the
  oracle's, with periodic folds, escapes and slot overflow;
* the reference program's tracking_method 1 through the drop-in shim, with
  and without trajectory diagnostics.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import golden
from test_particle_amd import abi
from test_particle_amd.fields import gyroperiod, mode_table
from test_particle_amd.particles import init_particles

pytestmark = pytest.mark.gpu

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
REF = os.path.join(ROOT, "oracle", "_ref")


@pytest.fixture(scope="module")
def gpu():
    from test_particle_amd.gpu import Tracker
    t = Tracker(0)
    yield t
    t.close()


def _same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    diff = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    assert not diff.any(), f"{what}: {diff.sum()} differ, first {np.argwhere(diff)[0]}"


def _compare(r, ref, what):
    for k in ("ptl", "nsteps", "espect", "espect_escape", "remain"):
        _same(r[k], ref[k], f"{what} {k}")
    for k in ("accepted", "rejected", "rhs", "failed"):
        assert r[k] == ref[k], f"{what} {k}: {r[k]} vs {ref[k]}"
    d, dr = r["dcoeff"], ref["dcoeff"]
    bad = np.abs(d - dr) > 1e-12 * np.abs(dr) + 1e-15
    assert not bad.any(), f"{what} D sums: first {np.argwhere(bad)[0]}"


def _case(name):
    if name == "uniform_periodic":
        ptl = init_particles(200, (0, 0, 0), (5, 5, 5), 0.6, seed=81)
        prm = abi.make_params(dt=0.05, estep=400, lo=(0, 0, 0), hi=(5, 5, 5), nbins=40,
                              nt_out=10, emin=1e-3, emax=1.0, einterval_t=40, dcoeff_interval=40)
        bs = abi.make_bs_params(rtol=1e-10, atol=1e-12, b_ref=1.0)
        return ("test", None), ptl, prm, bs
    rows = mode_table(64, kmin=0.3, kmax=30.0, rms=0.5, seed=82)
    ptl = init_particles(300, (0, 0, 0), (10, 10, 10), 0.1, seed=83, kind="loguniform",
                         vmax=0.9)
    bc = abi.TP_BC_OPEN if name == "modes_open" else abi.TP_BC_PERIODIC
    prm = abi.make_params(dt=0.05, estep=600, lo=(0, 0, 0), hi=(10, 10, 10), bc_flag=bc,
                          nbins=40, nt_out=12, emin=1e-3, emax=1.0, einterval_t=50,
                          dcoeff_interval=60)
    bs = abi.make_bs_params(rtol=1e-8, atol=1e-10, b_ref=1.0)
    return ("modes", rows), ptl, prm, bs


def _set(gpu, oracle, field):
    kind, rows = field
    if kind == "test":
        gpu.set_field_test()
        return oracle.Field.test()
    gpu.set_field_modes(rows, 1.0)
    return oracle.Field.modes(rows, 1.0)


@pytest.mark.parametrize("name", ["uniform_periodic", "modes_periodic", "modes_open"])
def test_dense_bitwise(gpu, oracle_tp, name):
    field, ptl, prm, bs = _case(name)
    f = _set(gpu, oracle_tp, field)
    print(f"[dense] {name}: {ptl.shape[0]} particles, {prm.estep} output steps", flush=True)
    r = gpu.track_adaptive_diag(bs, prm, ptl)
    ref = oracle_tp.bs_track_diag(f, bs, prm, ptl)
    _compare(r, ref, name)
    assert r["failed"] == 0
    if name == "modes_open":
        assert r["espect_escape"].sum() > 0 and r["remain"][-1] < ptl.shape[0]


def test_dense_launch_shapes(gpu, oracle_tp):
    field, ptl, prm, bs = _case("modes_periodic")
    f = _set(gpu, oracle_tp, field)
    ref = oracle_tp.bs_track_diag(f, bs, prm, ptl)
    for block, wpc in ((64, 1), (256, 2), (128, 8)):
        print(f"[dense] launch shape block {block} waves/CU {wpc}", flush=True)
        b = abi.make_bs_params(rtol=bs.rtol, atol=bs.atol, b_ref=1.0, block=block,
                               waves_per_cu=wpc)
        _compare(gpu.track_adaptive_diag(b, prm, ptl), ref, f"shape {block}/{wpc}")


def test_dense_pic_grid(gpu, oracle_tp):
    g = golden("track_grid")
    ptl = init_particles(150, g["lo"], g["hi"], 0.3, seed=84)
    prm = abi.make_params(dt=0.05, estep=200, lo=g["lo"], hi=g["hi"], nbins=30, nt_out=5,
                          emin=1e-3, emax=1.0, einterval_t=40, dcoeff_interval=50)
    bs = abi.make_bs_params(rtol=1e-8, atol=1e-10, b_ref=1.0)
    gpu.set_field_grid(g["egrid"], g["bgrid"], g["lo"], g["hi"])
    print("[dense] PIC grid", flush=True)
    r = gpu.track_adaptive_diag(bs, prm, ptl)
    ref = oracle_tp.bs_track_diag(oracle_tp.Field.grid(g["egrid"], g["bgrid"], g["lo"], g["hi"]),
                                  bs, prm, ptl)
    _compare(r, ref, "grid")


def _slots(n, nrec, seed):
    """nrec + 1 slots per particle, every third particle fewer (its
    later points are counted, not written); slot 0 the initial point, the
    rest a fill value the records must replace"""
    slots = np.full(n, nrec + 1)
    slots[::3] = np.random.default_rng(seed).integers(1, nrec + 1, slots[::3].size)
    accum = np.cumsum(slots).astype(np.int32)
    off = np.concatenate([[0], accum[:-1]])
    return slots, accum, off


@pytest.mark.parametrize("name", ["uniform_periodic", "modes_periodic", "modes_open"])
def test_dense_trajectory_bitwise(gpu, oracle_tp, name):
    field, ptl, prm, bs = _case(name)
    f = _set(gpu, oracle_tp, field)
    n, nout = ptl.shape[0], 30
    nrec = prm.estep // nout
    slots, accum, off = _slots(n, nrec, 85)
    pt = np.full((accum[-1], abi.NFIELD), -7.0)
    pt[off] = ptl
    print(f"[dense traj] {name}: {n} particles, output every {nout} steps", flush=True)
    r = gpu.track_adaptive_diag_traj(bs, prm, ptl, nout, accum, pt)
    ref = oracle_tp.bs_track_diag_traj(f, bs, prm, ptl, nout, accum, pt)
    _compare(r, ref, name)
    _same(r["ptl_time"], ref["ptl_time"], f"{name} ptl_time")
    _same(r["points"], ref["points"], f"{name} points")
    # the records replaced every fill value they reached
    for i in range(n):
        w = min(int(r["points"][i]), int(slots[i]))
        assert (r["ptl_time"][off[i]:off[i] + w, 7] != -7.0).all(), i
    if name == "modes_open":
        assert (r["points"] < nrec + 1).any()      # escaped particles stop recording
    else:
        assert (r["points"] == nrec + 1).all()
    # the diagnostics are those of the run without records
    plain = gpu.track_adaptive_diag(bs, prm, ptl)
    for k in ("ptl", "espect", "espect_escape", "remain"):
        _same(r[k], plain[k], f"{name} {k} with and without records")


def test_dense_trajectory_pic_grid(gpu, oracle_tp):
    """the records through the reference gather of the PIC grid, with
    overflowing slots"""
    g = golden("track_grid")
    ptl = init_particles(150, g["lo"], g["hi"], 0.3, seed=86)
    prm = abi.make_params(dt=0.05, estep=200, lo=g["lo"], hi=g["hi"], nbins=30, nt_out=5,
                          emin=1e-3, emax=1.0, einterval_t=40, dcoeff_interval=50)
    bs = abi.make_bs_params(rtol=1e-8, atol=1e-10, b_ref=1.0)
    gpu.set_field_grid(g["egrid"], g["bgrid"], g["lo"], g["hi"])
    f = oracle_tp.Field.grid(g["egrid"], g["bgrid"], g["lo"], g["hi"])
    n, nout = ptl.shape[0], 7
    slots, accum, off = _slots(n, prm.estep // nout, 87)
    pt = np.zeros((accum[-1], abi.NFIELD))
    pt[off] = ptl
    print("[dense traj] PIC grid", flush=True)
    r = gpu.track_adaptive_diag_traj(bs, prm, ptl, nout, accum, pt)
    ref = oracle_tp.bs_track_diag_traj(f, bs, prm, ptl, nout, accum, pt)
    _compare(r, ref, "grid traj")
    _same(r["ptl_time"], ref["ptl_time"], "grid ptl_time")
    _same(r["points"], ref["points"], "grid points")
    assert (r["points"] == prm.estep // nout + 1).all()


def test_dense_trajectory_arguments(gpu):
    """decreasing slot ranges and nsteps_output <= 0 are refused"""
    from test_particle_amd.gpu import TPError
    field, ptl, prm, bs = _case("uniform_periodic")
    gpu.set_field_test()
    ptl = ptl[:4]
    pt = np.zeros((8, abi.NFIELD))
    with pytest.raises(TPError):
        gpu.track_adaptive_diag_traj(bs, prm, ptl, 10, [2, 4, 3, 8], pt)
    with pytest.raises(TPError):
        gpu.track_adaptive_diag_traj(bs, prm, ptl, 0, [2, 4, 6, 8], pt)


def test_dense_vs_fixed_step_statistics(gpu):
    """The same particles, field and output grid through the plain FP64
    Boris kernel (dt = 0.01 T_g) and the adaptive stepper (rtol 1e-8):
    200 modes, 131,072 particles, 20 gyroperiods.  D_xx at every D row within
    3 standard errors of the difference (per-particle terms at the last row)
    and 2 %; spectra within 2 % (relative L1)."""
    rows = mode_table(200, kmin=0.05, kmax=5.0, rms=0.5, seed=85)
    n = 1 << 17
    L = 2 * np.pi / 0.05
    ptl = init_particles(n, (0, 0, 0), (L, L, L), 0.7, seed=86)
    dt = 0.01 * gyroperiod(0.7, 1.0)
    prm = abi.make_params(dt=dt, estep=2000, lo=(0, 0, 0), hi=(L, L, L), pusher=abi.TP_PUSH_BORIS,
                          nbins=60, nt_out=10, emin=1e-3, emax=1.0, dcoeff_interval=100)
    gpu.set_field_modes(rows, 1.0)
    print(f"[dense] statistics vs fixed step: {n} particles", flush=True)
    fx = gpu.track_fixed(prm, ptl)
    ad = gpu.track_adaptive_diag(abi.make_bs_params(rtol=1e-8, atol=1e-10, b_ref=1.0), prm, ptl)
    assert ad["failed"] == 0
    dxx = lambda r: r["dcoeff"][1, 1:] / r["remain"][1:]  # noqa: E731
    t_end = prm.estep * dt

    def terms(r):
        p = r["ptl"]
        return (ptl[:, 0] - (p[:, 0] + p[:, 8])) ** 2 / (2 * t_end + dt)
    a, b = terms(fx), terms(ad)
    se = np.hypot(a.std(ddof=1), b.std(ddof=1)) / np.sqrt(n)
    assert abs(a.mean() / dxx(fx)[-1] - 1) < 1e-9 and abs(b.mean() / dxx(ad)[-1] - 1) < 1e-9
    rel = np.abs(dxx(ad) / dxx(fx) - 1)
    print(f"D_xx rel diff per row: max {rel.max():.3e}, last {rel[-1]:.3e}; "
          f"last-row SE {se / a.mean():.3e}", flush=True)
    assert abs(b.mean() - a.mean()) <= 3 * se
    assert rel.max() < 0.02
    s = np.abs(ad["espect"] - fx["espect"]).sum(1) / fx["espect"].sum(1)
    assert s.max() < 0.02, s


def test_dropin_tracking_method_1(tmp_path, oracle_tp):
    """The reference program with tracking_method 1 (src/tracking.c:680-683,
    whose adaptive loop is commented out) through the GPU shim: the adaptive
    path runs (stderr line), with the reference's constants (KMAXX 8, atol
    1e-5, rtol 1e-14, first step 1e-5); its particles are the oracle's with
    those constants bit for bit, and data/espectrum.dat holds the spectra."""
    exe = os.path.join(REF, "ref_harness_gpu")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/ref_harness_gpu not built")
    lo, hi = (0.0, 0.0, 0.0), (5.0, 5.0, 5.0)
    ptl = init_particles(100, lo, hi, 0.5, seed=87)
    spec = dict(system=0, tmax=300 * 0.05 + 1e-9, dt=0.05, bc=0, nbins=30, nt_out=3,
                emin=1e-3, emax=1.0, method=1, xmin=lo[0], ymin=lo[1], zmin=lo[2],
                xmax=hi[0], ymax=hi[1], zmax=hi[2])
    os.makedirs(tmp_path / "data")
    with open(tmp_path / "spec.txt", "w") as f:
        for k, v in spec.items():
            f.write(f"{k} {v!r}\n" if isinstance(v, float) else f"{k} {v}\n")
    ptl.tofile(tmp_path / "p.bin")
    r = subprocess.run([exe, "hybrid", "spec.txt", "p.bin", "out.bin"], cwd=tmp_path,
                       timeout=120, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True,
                       env=dict(os.environ, OMP_NUM_THREADS="1", TP_SHIM_REQUIRE_GPU="1"))
    assert r.returncode == 0, r.stderr[-2000:]
    assert "particle_tracking_hybrid on the GPU" in r.stderr
    out = np.fromfile(tmp_path / "out.bin").reshape(-1, 11)
    prm = abi.make_params(dt=0.05, estep=300, lo=lo, hi=hi, nbins=30, nt_out=3, emin=1e-3,
                          emax=1.0)
    bs = abi.make_bs_params(rtol=1e-14, atol=1e-5, h0=1e-5, kmax=8, b_ref=1.0)
    ref = oracle_tp.bs_track_diag(oracle_tp.Field.test(), bs, prm, ptl)
    _same(out, ref["ptl"], "drop-in particles")
    lines = open(tmp_path / "data" / "espectrum.dat").read().splitlines()
    assert len(lines) == 2 + 3
    assert not os.path.exists(tmp_path / "data" / "diff_coeffs.dat")


def test_dropin_tracking_method_1_trajectories(tmp_path, oracle_tp):
    """traj_diagnose = 1 with tracking_method 1 through the drop-in shim
    (the call of src/diagnostics.c:555-557): the reference program's
    ptl_time and particles are the oracle's adaptive records bit for bit."""
    exe = os.path.join(REF, "ref_harness_gpu")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/ref_harness_gpu not built")
    lo, hi = (0.0, 0.0, 0.0), (5.0, 5.0, 5.0)
    n, estep, nout = 60, 300, 20
    ptl = init_particles(n, lo, hi, 0.5, seed=88)
    spec = dict(system=0, tmax=estep * 0.05 + 1e-9, dt=0.05, bc=0, nbins=30, nt_out=3,
                emin=1e-3, emax=1.0, method=1, nsteps_output=nout, xmin=lo[0], ymin=lo[1],
                zmin=lo[2], xmax=hi[0], ymax=hi[1], zmax=hi[2])
    os.makedirs(tmp_path / "o")
    with open(tmp_path / "spec.txt", "w") as f:
        for k, v in spec.items():
            f.write(f"{k} {v!r}\n" if isinstance(v, float) else f"{k} {v}\n")
    ptl.tofile(tmp_path / "p.bin")
    print("[dropin] tracking_method 1 trajh", flush=True)
    r = subprocess.run([exe, "trajh", "spec.txt", "p.bin", "o"], cwd=tmp_path, timeout=120,
                       stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True,
                       env=dict(os.environ, OMP_NUM_THREADS="1", TP_SHIM_REQUIRE_GPU="1"))
    assert r.returncode == 0, r.stderr[-2000:]
    assert "particle_tracking_hybrid on the GPU" in r.stderr
    slots = estep // nout + 1
    traj = np.fromfile(tmp_path / "o" / "traj.bin").reshape(-1, 11)
    out = np.fromfile(tmp_path / "o" / "ptl.bin").reshape(-1, 11)
    prm = abi.make_params(dt=0.05, estep=estep, lo=lo, hi=hi, nbins=30, nt_out=3, emin=1e-3,
                          emax=1.0)
    bs = abi.make_bs_params(rtol=1e-14, atol=1e-5, h0=1e-5, kmax=8, b_ref=1.0)
    accum = (np.arange(1, n + 1) * slots).astype(np.int32)
    pt = np.zeros((n * slots, 11))
    pt[::slots] = ptl
    ref = oracle_tp.bs_track_diag_traj(oracle_tp.Field.test(), bs, prm, ptl, nout, accum, pt)
    _same(out, ref["ptl"], "drop-in particles")
    _same(traj, ref["ptl_time"], "drop-in ptl_time")
    assert (ref["points"] == slots).all()
