"""GPU learned field surrogate (csrc/tp_mlp.hip) against the CPU oracle
(oracle/mlpref.c, itself pinned to PyTorch and finite differences in
test_oracle_mlp.py).

* forward: the GPU network (bf16 matrix cores, FP32 accumulation, hardware
  sin/cos/exp) against the oracle's emulation of the same rounding points;
  the only differences are the FP32 summation order and the hardware
  transcendentals, which can flip a bf16 rounding;
* gradients of the mean-squared loss, per parameter block;
* training on the GPU: the loss falls, the field residual against the
  exact mode sum falls, and two trainings from the same start are
  bitwise identical (fixed-order reductions);
* the surrogate field in the FP32 Boris step: with tolerance 0 every
  particle falls back at its first check and the run is bitwise the exact
  FP32 run; with the trained network fewer particles need the exact field
  than with the untrained one (what the training gains), and D_xx stays
  close to the exact run's.
"""
import numpy as np
import pytest

from test_particle_amd import abi, surrogate
from test_particle_amd.fields import gyroperiod
from test_particle_amd.particles import init_particles

pytestmark = pytest.mark.gpu

NMODES = 200


def case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def gpu():
    from test_particle_amd.gpu import Tracker
    t = Tracker(0)
    yield t
    t.close()


@pytest.fixture(scope="module")
def net():
    rows, period = surrogate.surrogate_field(NMODES)
    q = surrogate.fourier_frequencies(period, sigma=3.0, seed=3)
    p0 = surrogate.init_params(3)
    scale = surrogate.output_scale(rows)
    return dict(rows=rows, period=period, q=q, p0=p0, scale=scale)


@pytest.fixture(scope="module")
def trained(gpu, net):
    case("train surrogate: 2^20 samples, batch 16384, 3 epochs")
    gpu.set_field_modes(net["rows"], 1.0)
    gpu.mlp_set(net["q"], net["p0"], net["scale"])
    prm = abi.make_mlp_train_params(nsamples=1 << 20, period=net["period"], batch=16384,
                                    epochs=3, lr=2e-3, lr_min=2e-5, seed=11)
    st = gpu.train_surrogate(prm)
    return dict(stats=st, params=gpu.mlp_get(), prm=prm)


def test_forward_matches_oracle(gpu, net):
    from oracle import tpref
    case("mlp forward, 1000 points")
    gpu.set_field_modes(net["rows"], 1.0)
    gpu.mlp_set(net["q"], net["p0"], net["scale"])
    rng = np.random.default_rng(1)
    x = rng.uniform(0.0, net["period"], (1000, 3))
    out = gpu.mlp_forward(x)
    p64 = net["p0"].astype(np.float64)
    q64 = net["q"].astype(np.float64)
    ref = tpref.mlp_forward(p64, q64, x, emulate=True)
    scale = np.max(np.abs(ref))
    err = np.abs(out - ref)
    assert np.max(err) < 3e-3 * scale, np.max(err) / scale
    assert np.sqrt(np.mean(err ** 2)) < 3e-4 * scale
    full = tpref.mlp_forward(p64, q64, x)
    assert np.max(np.abs(out - full)) < 3e-2 * np.max(np.abs(full))


def test_loss_grad_matches_oracle(gpu, net):
    from oracle import tpref
    case("mlp loss and gradient, 300 samples")
    gpu.mlp_set(net["q"], net["p0"], net["scale"])
    rng = np.random.default_rng(2)
    x = rng.uniform(0.0, net["period"], (300, 3)).astype(np.float32)
    y = rng.normal(0.0, 1.0, (300, 3)).astype(np.float32)
    loss, g = gpu.mlp_loss_grad(x, y)
    lref, gref = tpref.mlp_loss_grad(net["p0"].astype(np.float64), net["q"].astype(np.float64),
                                     x.astype(np.float64), y.astype(np.float64), emulate=True)
    assert abs(loss - lref) < 1e-4 * lref
    o = 0
    for fin, fout in surrogate.LAYERS:
        for size, what in ((fin * fout, f"W{fin}x{fout}"), (fout, f"b{fout}")):
            a, b = g[o:o + size].astype(np.float64), gref[o:o + size]
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            assert rel < 1e-2, (what, o, rel)
            o += size
    assert o == abi.TP_MLP_NPARAMS


def _residual(gpu, net, x):
    pts = np.zeros((x.shape[0], 4))
    pts[:, :3] = x
    exact = gpu.eval_field(pts)[:, :3]
    b = net["scale"] * gpu.mlp_forward(x)
    b[:, 2] += 1.0
    return np.linalg.norm(b - exact, axis=1) / np.linalg.norm(exact, axis=1)


def test_training_reduces_loss_and_residual(gpu, net, trained):
    case("trained surrogate residual, 8192 points")
    st = trained["stats"]
    assert st["steps"] == 3 * (1 << 20) // 16384
    assert st["last_loss"] < 0.05 * st["first_loss"], st
    x = np.random.default_rng(4).uniform(0.0, net["period"], (8192, 3))
    gpu.set_field_modes(net["rows"], 1.0)
    gpu.mlp_set(net["q"], net["p0"], net["scale"])
    r0 = _residual(gpu, net, x)
    gpu.mlp_set(net["q"], trained["params"], net["scale"])
    r1 = _residual(gpu, net, x)
    print(f"residual median untrained {np.median(r0):.3e}, trained {np.median(r1):.3e}")
    assert np.median(r1) < 0.2 * np.median(r0)


def test_training_is_deterministic(gpu, net, trained):
    case("retrain from the same start")
    gpu.set_field_modes(net["rows"], 1.0)
    gpu.mlp_set(net["q"], net["p0"], net["scale"])
    st = gpu.train_surrogate(trained["prm"])
    assert np.array_equal(gpu.mlp_get(), trained["params"])
    assert st["last_loss"] == trained["stats"]["last_loss"]


def _run(gpu, net, ptl, steps, mlp=None):
    period = net["period"]
    lo, hi = (0.0, 0.0, 0.0), (period,) * 3
    prm = abi.make_params(dt=0.01 * gyroperiod(0.7, 1.0), estep=steps, lo=lo, hi=hi,
                          pusher=abi.TP_PUSH_BORIS, precision=abi.TP_FP32_MIXED, nbins=102,
                          nt_out=max(1, steps // 100), emin=1e-4, emax=1.0)
    gpu.set_field_modes(net["rows"], 1.0)
    nfb = None
    if mlp is not None:
        params, tol = mlp
        gpu.mlp_set(net["q"], params, net["scale"])
        gpu.set_field_mlp(tol, 50)
    r = gpu.track_fixed(prm, ptl)
    if mlp is not None:
        nfb = gpu.mlp_fallback_count()
    return r, nfb


def test_tolerance_zero_is_the_exact_run(gpu, net, trained):
    n, steps = 3000, 300
    case(f"surrogate field, tol 0: {n} particles x {steps} steps vs exact FP32")
    ptl = init_particles(n, (0, 0, 0), (net["period"],) * 3, 0.7, seed=5)
    ref, _ = _run(gpu, net, ptl, steps)
    r, nfb = _run(gpu, net, ptl, steps, mlp=(trained["params"], 0.0))
    assert nfb == n
    assert np.array_equal(r["ptl"], ref["ptl"])
    assert np.array_equal(r["espect"], ref["espect"])
    assert np.array_equal(r["nsteps"], ref["nsteps"])
    assert np.allclose(r["dcoeff"], ref["dcoeff"], rtol=1e-12, atol=0)


def test_trained_network_saves_exact_evaluations(gpu, net, trained):
    n, steps = 8192, 500
    ptl = init_particles(n, (0, 0, 0), (net["period"],) * 3, 0.7, seed=6)
    x = np.random.default_rng(4).uniform(0.0, net["period"], (8192, 3))
    gpu.set_field_modes(net["rows"], 1.0)
    gpu.mlp_set(net["q"], trained["params"], net["scale"])
    tol = float(np.quantile(_residual(gpu, net, x), 0.9))
    case(f"surrogate field, tol {tol:.3e}: {n} particles x {steps} steps")
    ref, _ = _run(gpu, net, ptl, steps)
    r_tr, nfb_tr = _run(gpu, net, ptl, steps, mlp=(trained["params"], tol))
    r_un, nfb_un = _run(gpu, net, ptl, steps, mlp=(net["p0"], tol))
    print(f"fallbacks: trained {nfb_tr}, untrained {nfb_un} of {n}")
    assert nfb_un > 0.95 * n
    assert nfb_tr < 0.5 * n
    # every particle advanced once per step, whichever list it was in
    assert np.array_equal(r_tr["nsteps"], ref["nsteps"])
    assert np.array_equal(r_un["nsteps"], ref["nsteps"])
    dxx = lambda r: r["dcoeff"][1, 1:] / np.maximum(r["remain"][1:], 1)  # noqa: E731
    err = np.max(np.abs(dxx(r_tr) / dxx(ref) - 1.0))
    print(f"D_xx max rel err vs exact: trained {err:.3e}")
    assert err < 0.05


def test_network_field_drives_the_push(gpu, net, trained):
    """One step with no fallback (tolerance 1e30): the velocity change is
    the Boris rotation in the network's field at the start positions (the
    field from tp_gpu_mlp_forward, the push redone in FP64 on the host)."""
    n = 1000
    case(f"network-field push, {n} particles x 1 step")
    ptl = init_particles(n, (0, 0, 0), (net["period"],) * 3, 0.7, seed=8)
    r, nfb = _run(gpu, net, ptl, 1, mlp=(trained["params"], 1e30))
    assert nfb == 0
    b = net["scale"] * gpu.mlp_forward(ptl[:, :3])
    b[:, 2] += 1.0
    dt = 0.01 * gyroperiod(0.7, 1.0)
    u = ptl[:, 4:7].astype(np.float32).astype(np.float64)
    h = 0.5 * dt
    gm = np.sqrt(1.0 + np.sum(u * u, axis=1, keepdims=True))
    t = h * b / gm
    sf = 2.0 / (1.0 + np.sum(t * t, axis=1, keepdims=True))
    up = u + np.cross(u, t)
    uq = u + np.cross(up, sf * t)
    assert np.max(np.abs(r["ptl"][:, 4:7] - uq)) < 1e-5 * np.max(np.abs(uq))
    g = np.sqrt(1.0 + np.sum(uq * uq, axis=1, keepdims=True))
    x1 = ptl[:, :3] + dt * uq / g
    x1 = np.mod(x1, net["period"])
    d = np.abs(r["ptl"][:, :3] - x1)
    d = np.minimum(d, net["period"] - d)
    assert np.max(d) < 1e-5 * dt
