"""The learned-surrogate workload at its configured size, on the GPU.

1000-mode turbulence (the adaptive workload's modes on the lattice of its
own k_min, period L = 2 pi / k_min), the network trained on the GPU on
2^24 exact FP64 samples (4 epochs), then 4,194,304 particles at |v| = 0.7c,
5,000 FP32 Boris steps, dt = 0.01 T_g, periodic in L, residual checks every
100 steps.  Against the exact-field FP32 run on the same particles
(the field slot the network fills: get_emf, src/emfields.c:32-60):

* tolerance 0: every particle falls back at the first check and the run is
  bitwise the exact run (particles, step counts, spectra; D sums 1e-12);
* tolerance 1e-3 (the configured one) and the residual's 90th percentile
  at uniform points (the network in use: a particle falls back at the first
  of the 50 checks its residual fails, and the network carries at least a
  quarter of the particles through all 5,000 steps): every particle's step
  count is the exact run's, D_xx and the energy spectra within 1 %;
* the first residual check: on 512 sampled particles the fallback flags the
  kernel set agree with the residual recomputed from tp_gpu_mlp_forward and
  the FP32 exact field of tp_gpu_eval_field at the same positions (apart
  from residuals within 1e-4 of the tolerance, where the FP32 arithmetic of
  the two evaluations may round either way)."""
import numpy as np
import pytest

from test_particle_amd import abi, surrogate
from test_particle_amd.fields import gyroperiod
from test_particle_amd.particles import init_particles

pytestmark = pytest.mark.gpu

N = 1 << 22
STEPS = 5000
SEED = 20261202


def case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def setup():
    from test_particle_amd.gpu import Tracker
    rows, period = surrogate.surrogate_field(1000, seed=SEED)
    t = Tracker(0)
    case("train the surrogate: 2^24 samples, batch 16384, 4 epochs")
    t.set_field_modes(rows, 1.0)
    q = surrogate.fourier_frequencies(period, sigma=4.0, seed=SEED)
    scale = surrogate.output_scale(rows)
    t.mlp_set(q, surrogate.init_params(SEED), scale)
    st = t.train_surrogate(abi.make_mlp_train_params(
        nsamples=1 << 24, period=period, batch=16384, epochs=4, lr=2e-3, lr_min=2e-5,
        seed=SEED))
    lo, hi = (0.0, 0.0, 0.0), (period,) * 3
    ptl = init_particles(N, lo, hi, 0.7, seed=SEED)
    prm = abi.make_params(dt=0.01 * gyroperiod(0.7, 1.0), estep=STEPS, lo=lo, hi=hi,
                          pusher=abi.TP_PUSH_BORIS, precision=abi.TP_FP32_MIXED, nbins=102,
                          nt_out=100, emin=1e-4, emax=1.0)
    case(f"exact FP32 run: {N} particles x {STEPS} steps")
    t.set_field_modes(rows, 1.0)
    exact = t.track_fixed(prm, ptl)
    yield dict(t=t, rows=rows, period=period, scale=scale, ptl=ptl, prm=prm, exact=exact,
               train=st)
    t.close()


def _net_run(s, tol, params=None):
    t = s["t"]
    t.set_field_mlp(tol, 100)
    r = t.track_fixed(params or s["prm"], s["ptl"])
    flags = t.mlp_fallback_flags()
    t.set_field_modes(s["rows"], 1.0)
    return r, flags


def _dxx(r):
    return r["dcoeff"][1, 1:] / np.maximum(r["remain"][1:], 1)


def _spectrum_err(a, b):
    return float(np.max(np.abs(a - b).sum(axis=1) / np.maximum(b.sum(axis=1), 1.0)))


def _residual(s, x):
    t = s["t"]
    pts = np.zeros((x.shape[0], 4))
    pts[:, :3] = x
    exact = t.eval_field(pts, abi.TP_FP32_MIXED)[:, :3]
    b = s["scale"] * t.mlp_forward(x)
    b[:, 2] += 1.0
    return np.linalg.norm(b - exact, axis=1) / np.linalg.norm(exact, axis=1)


def test_training_converged(setup):
    st = setup["train"]
    assert st["steps"] == 4 * (1 << 24) // 16384
    assert st["last_loss"] < 1e-2 * st["first_loss"], st


def test_tolerance_zero_is_the_exact_run(setup):
    case("network field, tol 0, full workload")
    r, flags = _net_run(setup, 0.0)
    ex = setup["exact"]
    assert flags.all()
    assert np.array_equal(r["ptl"], ex["ptl"])
    assert np.array_equal(r["nsteps"], ex["nsteps"])
    assert np.array_equal(r["espect"], ex["espect"])
    assert np.allclose(r["dcoeff"], ex["dcoeff"], rtol=1e-12, atol=0.0)


@pytest.mark.parametrize("which", ["configured_1e-3", "p90_residual"])
def test_full_workload_within_one_percent(setup, which):
    s = setup
    if which == "configured_1e-3":
        tol = 1e-3
    else:
        x = np.random.default_rng(SEED + 1).uniform(0.0, s["period"], (1 << 16, 3))
        tol = float(np.quantile(_residual(s, x), 0.9))
    case(f"network field, tol {tol:.3e}, full workload")
    r, flags = _net_run(s, tol)
    ex = s["exact"]
    frac = float(flags.mean())
    derr = float(np.max(np.abs(_dxx(r) / _dxx(ex) - 1.0)))
    serr = _spectrum_err(r["espect"], ex["espect"])
    print(f"tol {tol:.3e}: fallback fraction {frac:.4f}, D_xx max rel err {derr:.2e}, "
          f"spectrum L1 err {serr:.2e}", flush=True)
    assert np.array_equal(r["nsteps"], ex["nsteps"])
    assert derr <= 0.01
    assert serr <= 0.01
    if which == "p90_residual":
        assert frac < 0.75           # the network carries a quarter or more


def test_first_check_flags_match_residuals(setup):
    s = setup
    x = np.random.default_rng(SEED + 1).uniform(0.0, s["period"], (1 << 16, 3))
    tol = float(np.quantile(_residual(s, x), 0.5))
    case(f"first residual check, tol {tol:.3e}: flags of 512 sampled particles")
    one = abi.make_params(dt=s["prm"].dt, estep=1, lo=(0.0, 0.0, 0.0), hi=(s["period"],) * 3,
                          pusher=abi.TP_PUSH_BORIS, precision=abi.TP_FP32_MIXED, nbins=102,
                          nt_out=1, emin=1e-4, emax=1.0)
    _, flags = _net_run(s, tol, one)
    idx = np.random.default_rng(7).choice(N, 512, replace=False)
    res = _residual(s, s["ptl"][idx, :3].astype(np.float32).astype(np.float64))
    expect = res > tol
    clear = np.abs(res / tol - 1.0) > 1e-4
    assert 0.2 < expect.mean() < 0.8
    assert np.array_equal(flags[idx][clear].astype(bool), expect[clear])
