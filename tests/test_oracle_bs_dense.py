"""The oracle's adaptive stepper with dense output and the fixed-step loop's
diagnostics (tpref_bs_track_diag), pinned by closed forms and scipy.

The reference's dense output (prepare_dense / dense_interp / dense_out,
src/unused/StepperBS.c:478-644) never ran: its caller is commented out
(src/tracking.c:518-630) and the file does not compile.  The restatement is
pinned here instead:

* uniform B (the reference's system 0, src/emfields.c:210-218): the exact
  helix gives every particle's position and momentum at every output time,
  so the D rows (calc_diff_coeff's terms summed, src/diagnostics.c:659-705)
  follow in closed form -- including periodic wraps, whose shifts the terms
  must undo; the energy is conserved, so every frame of the spectrum is the
  initial histogram;
* a multi-mode field: the D rows against scipy's DOP853 dense solution of
  the same ODE at the same output times;
* open boundaries: a particle streaming along B leaves at a known time;
  frames before it count the particle, the escape spectrum and the
  remaining counts account for it once.
"""
import numpy as np

from test_particle_amd import abi
from test_particle_amd.fields import eval_modes_numpy, mode_table
from test_particle_amd.particles import init_particles


def _helix(p, t):
    """Exact state at time t in B = z-hat (q/m = 1): u rotates at 1/gamma."""
    g = np.sqrt(1 + (p[:, 4:7] ** 2).sum(1))
    ph = t / g
    ux0, uy0, uz0 = p[:, 4], p[:, 5], p[:, 6]
    ux = ux0 * np.cos(ph) + uy0 * np.sin(ph)
    uy = -ux0 * np.sin(ph) + uy0 * np.cos(ph)
    x = p[:, 0] + ux0 * np.sin(ph) + uy0 * (1 - np.cos(ph))
    y = p[:, 1] + ux0 * (np.cos(ph) - 1) + uy0 * np.sin(ph)
    z = p[:, 2] + uz0 / g * t
    return np.column_stack([x, y, z]), np.column_stack([ux, uy, uz0])


def test_dense_output_follows_the_helix(oracle_tp):
    n, dt, estep, ei, di = 32, 0.05, 400, 40, 40
    # periodic wraps (particle_bc wraps once per axis and accepted step, so
    # the box is larger than a step's displacement)
    box = ((0.0, 0.0, 0.0), (5.0, 5.0, 5.0))
    p = init_particles(n, *box, 0.6, seed=71)
    prm = abi.make_params(dt=dt, estep=estep, lo=box[0], hi=box[1], nbins=40, nt_out=10,
                          emin=1e-3, emax=1.0, einterval_t=ei, dcoeff_interval=di)
    bs = abi.make_bs_params(rtol=1e-10, atol=1e-12, b_ref=1.0)
    r = oracle_tp.bs_track_diag(oracle_tp.Field.test(), bs, prm, p)
    assert r["failed"] == 0 and r["rejected"] >= 0 and r["accepted"] > n
    nd = estep // di + 1
    for k in range(1, nd):
        t = k * di * dt
        x, u = _helix(p, t)
        den = 2 * t + dt
        dxx = ((p[:, 0] - x[:, 0]) ** 2 / den).sum()
        dyy = ((p[:, 1] - x[:, 1]) ** 2 / den).sum()
        dzz = ((p[:, 2] - x[:, 2]) ** 2 / den).sum()
        dpp = (((p[:, 4:7] - u) ** 2).sum(1) / den).sum()
        got = r["dcoeff"][:, k]
        assert abs(got[1] / dxx - 1) < 1e-7, (k, got[1], dxx)
        assert abs(got[2] / dyy - 1) < 1e-7, (k, got[2], dyy)
        assert abs(got[3] / dzz - 1) < 1e-7, (k, got[3], dzz)
        assert abs(got[0] / (dxx + dyy + dzz) - 1) < 1e-7
        assert abs(got[4] / dpp - 1) < 1e-6
        assert abs(got[7]) < 1e-16 * got[1]             # energy is conserved
    assert (r["dcoeff"][:, 0] == 0).all()
    assert (r["remain"] == n).all()
    # every frame is the initial histogram
    e0 = np.sqrt(1 + (p[:, 4:7] ** 2).sum(1)) - 1
    b0 = np.clip(np.floor((np.log10(e0) - prm.logemin) / prm.logde), 0, prm.nbins - 1)
    h0 = np.bincount(b0.astype(int), minlength=prm.nbins)
    assert (r["espect"] == h0[None, :]).all()
    assert r["espect_escape"].sum() == 0
    # the final state is the helix at the end time, wrapped into the box
    x, u = _helix(p, estep * dt)
    assert np.abs(r["ptl"][:, 7] - estep * dt).max() < 1e-12
    assert np.abs(r["ptl"][:, :3] + r["ptl"][:, 8:11] - x).max() < 1e-8
    assert np.abs(r["ptl"][:, 4:7] - u).max() < 1e-8
    assert ((r["ptl"][:, :3] >= 0) & (r["ptl"][:, :3] <= 5)).all()
    assert (r["ptl"][:, 8:11] != 0).any()            # some particles wrapped


def test_dense_output_against_dop853(oracle_tp):
    from scipy.integrate import solve_ivp
    rows = mode_table(50, kmin=0.5, kmax=5.0, rms=0.5, seed=72)
    f = oracle_tp.Field.modes(rows, 1.0)
    n, dt, estep, di = 6, 0.1, 30, 5
    big = ((-1e3, -1e3, -1e3), (1e3, 1e3, 1e3))      # no wraps
    p = init_particles(n, (0, 0, 0), (3, 3, 3), 0.5, seed=73)
    prm = abi.make_params(dt=dt, estep=estep, lo=big[0], hi=big[1], nbins=20, nt_out=10,
                          emin=1e-3, emax=1.0, einterval_t=3, dcoeff_interval=di)
    bs = abi.make_bs_params(rtol=1e-10, atol=1e-12, b_ref=1.0)
    r = oracle_tp.bs_track_diag(f, bs, prm, p)

    def rhs(t, y):
        b = eval_modes_numpy(rows, y[None, :3], 1.0)[0]
        g = np.sqrt(1 + y[3:] @ y[3:])
        v = y[3:] / g
        return np.concatenate([v, np.cross(v, b)])

    times = np.arange(1, estep // di + 1) * di * dt
    xs = np.zeros((len(times), n, 6))
    for i in range(n):
        sol = solve_ivp(rhs, (0, times[-1]), np.concatenate([p[i, :3], p[i, 4:7]]),
                        method="DOP853", rtol=1e-12, atol=1e-14, dense_output=True)
        xs[:, i, :] = sol.sol(times).T
    for k, t in enumerate(times, start=1):
        den = 2 * t + dt
        dxx = ((p[:, 0] - xs[k - 1, :, 0]) ** 2 / den).sum()
        dpp = (((p[:, 4:7] - xs[k - 1, :, 3:]) ** 2).sum(1) / den).sum()
        assert abs(r["dcoeff"][1, k] / dxx - 1) < 1e-6, (k, r["dcoeff"][1, k], dxx)
        assert abs(r["dcoeff"][4, k] / dpp - 1) < 1e-5, (k, r["dcoeff"][4, k], dpp)
    assert r["failed"] == 0


def test_dense_output_open_boundary_accounting(oracle_tp):
    """Particles gyrating along B = z-hat with a drift reach z = 10 at
    t_e = (10 - z0) gamma / uz.  The escape is seen at the end of the
    accepted step, t_d in [t_e, t_e + H] (H bounds a step: the gyration
    limits it); half the particles have no drift and never leave.  Each
    escape is binned once; a frame at t_f counts the particles with
    t_d > t_f; row k counts out those with t_d <= t_(k-1), one row late, as
    the fixed-step loop does (src/tracking.c:366-369)."""
    n, dt, estep, ei, di = 40, 0.1, 300, 30, 30
    H = 3.0
    rng = np.random.default_rng(74)
    p = np.zeros((n, 11))
    p[:, 0:2] = rng.uniform(3, 7, (n, 2))
    p[:, 2] = rng.uniform(0.5, 9.5, n)
    p[:, 3] = 1.0
    p[:, 4] = 0.3
    p[:, 6] = np.where(np.arange(n) < n // 2, 0.4, 0.0)
    prm = abi.make_params(dt=dt, estep=estep, lo=(0, 0, 0), hi=(10, 10, 10), bc_flag=1,
                          nbins=20, nt_out=10, emin=1e-4, emax=1.0, einterval_t=ei,
                          dcoeff_interval=di)
    bs = abi.make_bs_params(rtol=1e-9, atol=1e-12, b_ref=1.0)
    r = oracle_tp.bs_track_diag(oracle_tp.Field.test(), bs, prm, p)
    g = np.sqrt(1 + (p[:, 4:7] ** 2).sum(1))
    vz = np.where(p[:, 6] > 0, p[:, 6], 1.0) / g
    te = np.where(p[:, 6] > 0, (10 - p[:, 2]) / vz, np.inf)
    tend = estep * dt
    assert r["espect_escape"].sum() == (te < tend).sum()
    assert r["espect_escape"].sum() == r["ptl"][:, 2].__gt__(10).sum()
    for f in range(10):
        tf = (f + 1) * ei * dt
        cnt = r["espect"][f].sum()
        assert (te > tf).sum() <= cnt <= (te > tf - H).sum(), (f, cnt)
    for k in range(1, estep // di + 1):
        tk1 = (k - 1) * di * dt
        out = n - r["remain"][k]
        assert (te + H <= tk1).sum() <= out <= (te <= tk1).sum(), (k, out)
    assert r["remain"][0] == n
    assert (np.diff(r["remain"]) <= 0).all()


def test_dense_trajectory_records_follow_the_helix(oracle_tp):
    """Trajectory diagnostics of the adaptive loop: the particle at the
    output steps j = k nsteps_output, from the dense output inside a step and
    the final state at the end time, folded into the periodic box as the
    fixed-step loop's records are (position + shift = the unwrapped helix).
    Slot 0 keeps the caller's point; points past a particle's slots are
    counted, not written."""
    n, dt, estep, nout = 24, 0.05, 400, 25
    box = ((0.0, 0.0, 0.0), (5.0, 5.0, 5.0))
    p = init_particles(n, *box, 0.6, seed=72)
    prm = abi.make_params(dt=dt, estep=estep, lo=box[0], hi=box[1], nbins=40, nt_out=10,
                          emin=1e-3, emax=1.0, einterval_t=40, dcoeff_interval=40)
    bs = abi.make_bs_params(rtol=1e-10, atol=1e-12, b_ref=1.0)
    nrec = estep // nout                             # 16 records after slot 0
    slots = np.full(n, nrec + 1)
    slots[::3] = 5                                   # these overflow
    accum = np.cumsum(slots).astype(np.int32)
    pt = np.full((accum[-1], 11), -7.0)
    off = np.concatenate([[0], accum[:-1]])
    pt[off] = p
    fld = oracle_tp.Field.test()
    r = oracle_tp.bs_track_diag_traj(fld, bs, prm, p, nout, accum, pt)
    plain = oracle_tp.bs_track_diag(fld, bs, prm, p)
    # the diagnostics are those of the run without records
    for k in ("ptl", "espect", "dcoeff", "remain", "nsteps"):
        assert np.array_equal(r[k], plain[k]), k
    assert (r["points"] == nrec + 1).all()
    tp = r["ptl_time"]
    assert np.array_equal(tp[off], p)
    for i in range(n):
        got = tp[off[i] + 1: off[i] + slots[i]]
        k = np.arange(1, got.shape[0] + 1)
        t = k * nout * dt
        x, u = _helix(np.repeat(p[i:i + 1], k.size, axis=0), t)
        assert np.abs(got[:, 7] - t).max() < 1e-12
        assert np.abs(got[:, 3] - p[i, 3]).max() == 0.0
        assert np.abs(got[:, :3] + got[:, 8:11] - x).max() < 1e-8, i
        assert np.abs(got[:, 4:7] - u).max() < 1e-8, i
        assert (got[:, :3] >= 0.0).all() and (got[:, :3] <= 5.0).all()
    # the last record is the final state
    full = np.flatnonzero(slots == nrec + 1)
    assert np.allclose(tp[off[full] + nrec][:, :7], r["ptl"][full][:, :7], rtol=0, atol=0)


def test_dense_trajectory_records_stop_at_the_escape(oracle_tp):
    """Open boundaries: a particle is recorded at the output steps before
    and inside the step in which it leaves, and never after."""
    n, dt, estep, nout = 16, 0.1, 300, 10
    rng = np.random.default_rng(75)
    p = np.zeros((n, 11))
    p[:, 0:2] = rng.uniform(3, 7, (n, 2))
    p[:, 2] = rng.uniform(0.5, 9.5, n)
    p[:, 3] = 1.0
    p[:, 4] = 0.3
    p[:, 6] = np.where(np.arange(n) < n // 2, 0.4, 0.0)
    prm = abi.make_params(dt=dt, estep=estep, lo=(0, 0, 0), hi=(10, 10, 10), bc_flag=1,
                          nbins=20, nt_out=10, emin=1e-4, emax=1.0, einterval_t=30,
                          dcoeff_interval=30)
    bs = abi.make_bs_params(rtol=1e-9, atol=1e-12, b_ref=1.0)
    accum = np.cumsum(np.full(n, estep // nout + 1)).astype(np.int32)
    pt = np.zeros((accum[-1], 11))
    r = oracle_tp.bs_track_diag_traj(oracle_tp.Field.test(), bs, prm, p, nout, accum, pt)
    g = np.sqrt(1 + (p[:, 4:7] ** 2).sum(1))
    te = np.where(p[:, 6] > 0, (10 - p[:, 2]) / (np.maximum(p[:, 6], 1e-300) / g), np.inf)
    H = 3.0                                          # bounds an accepted step
    for i in range(n):
        m = r["points"][i] - 1                       # records written
        if np.isinf(te[i]):
            assert m == estep // nout, i
        else:
            # records at t = k nout dt up to the end of the escaping step
            lo_k = int(np.floor(te[i] / (nout * dt)))
            hi_k = int(np.floor((te[i] + H) / (nout * dt)))
            assert lo_k <= m <= hi_k, (i, m, lo_k, hi_k)
