"""D_xx of the configured ABC workload on the GPU against the reference's
own tracking loop, compared statistically over three seeds.

Workload: 16,384 protons at |v| = 0.7c, ABC field A = B = C = 1, periodic
[0, 2 pi]^3, dt = 0.01 T_g (T_g = 2 pi gamma0 / sqrt(3)), 10,000 FP64 Wirz
steps (tracking_wirz, src/tracking.c:749-892), D every 100 steps
(calc_diff_coeff, src/diagnostics.c:659-705, summed at src/tracking.c:347-365).

The chain of comparisons, per seed:
* the reference: particle_tracking_hybrid compiled from its sources with
  A = B = C = 1 (oracle/_ref/ref_harness_abc1, mode hybrid, OpenMP) -- its
  final particles;
* the oracle with libm sin/cos, the reference's arithmetic: its final
  particles are bitwise the reference's over the full 10,000 steps;
* the GPU (bitwise the oracle with the project's sin/cos, tested in
  test_gpu_fullsize.py): its trajectories differ from the reference's only
  through the ulps of sin/cos, which the chaotic flow amplifies until the two
  runs are independent samples of the same diffusion.

So D_xx is compared statistically.  With dxx_i = (x0_i - (x_i + xshift_i))^2
/ (2 t_i + dt) the per-particle term at the last step, D = mean(dxx_i) and
SE = std(dxx_i) / sqrt(N) its standard error; the test asserts
|D_gpu - D_ref| <= K_SE * sqrt(SE_gpu^2 + SE_ref^2) with K_SE = 3 for every
seed, and prints the seed-to-seed spread of D_ref beside the mean SE."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from test_particle_amd import abi
from test_particle_amd.fields import abc_b_ref, gyroperiod
from test_particle_amd.particles import init_particles

pytestmark = pytest.mark.gpu

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
EXE = os.path.join(ROOT, "oracle", "_ref", "ref_harness_abc1")
SEEDS = (20261202, 20261203, 20261204)
N, STEPS = 16384, 10000
K_SE = 3.0
TWO_PI = 2.0 * np.pi
DT = 0.01 * gyroperiod(0.7, abc_b_ref())


def _cores():
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(env)) if env and env.isdigit() else n)


def _params():
    return abi.make_params(dt=DT, estep=STEPS, lo=(0.0,) * 3, hi=(TWO_PI,) * 3,
                           pusher=abi.TP_PUSH_WIRZ, nbins=102, nt_out=100, emin=1e-3,
                           emax=10.0)


def _reference(ptl):
    with tempfile.TemporaryDirectory() as tmp:
        os.makedirs(os.path.join(tmp, "data"))
        with open(os.path.join(tmp, "spec.txt"), "w") as f:
            f.write(f"system 2\nnx 5\nny 5\nnz 5\ntmax {float(STEPS * DT + 1e-9)!r}\n"
                    f"dt {float(DT)!r}\nbc 0\nnbins 102\nnt_out 100\nemin 0.001\n"
                    f"emax 10.0\n")
            for k in ("x", "y", "z"):
                f.write(f"{k}min 0.0\n{k}max {TWO_PI!r}\n")
        ptl.tofile(os.path.join(tmp, "p.bin"))
        subprocess.run([EXE, "hybrid", "spec.txt", "p.bin", "out.bin"], cwd=tmp, check=True,
                       capture_output=True, timeout=600,
                       env=dict(os.environ, OMP_NUM_THREADS=str(_cores())))
        out = np.fromfile(os.path.join(tmp, "out.bin")).reshape(-1, abi.NFIELD)
        diff = np.loadtxt(os.path.join(tmp, "data", "diff_coeffs.dat"), skiprows=1)
    return out, diff


def _dxx_terms(p0, p):
    return (p0[:, 0] - (p[:, 0] + p[:, 8])) ** 2 / (2.0 * p[:, 7] + DT)


@pytest.fixture(scope="module")
def runs():
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/ref_harness_abc1 not built")
    from oracle import tpref
    from test_particle_amd.gpu import Tracker
    out = {}
    with Tracker(0) as t:
        t.set_field_abc(1.0, 1.0, 1.0)
        for seed in SEEDS:
            print(f"[dxx] seed {seed}: GPU, reference, libm oracle", flush=True)
            ptl = init_particles(N, (0.0,) * 3, (TWO_PI,) * 3, 0.7, seed=seed)
            g = t.track_fixed(_params(), ptl)
            ref, diff = _reference(ptl)
            tpref.set_math(True)
            orc = tpref.track_fixed(tpref.Field.abc(1.0, 1.0, 1.0), _params(), ptl,
                                    nthreads=_cores())
            out[seed] = dict(ptl=ptl, gpu=g, ref=ref, ref_diff=diff, libm=orc)
    return out


def test_libm_oracle_is_the_reference_over_all_steps(runs):
    for seed, r in runs.items():
        assert np.array_equal(r["libm"]["ptl"], r["ref"]), seed


def test_dxx_gpu_vs_reference_within_k_standard_errors(runs):
    d_ref = []
    ses = []
    for seed, r in runs.items():
        a = _dxx_terms(r["ptl"], r["gpu"]["ptl"])
        b = _dxx_terms(r["ptl"], r["ref"])
        da, db = a.mean(), b.mean()
        sa, sb = a.std(ddof=1) / np.sqrt(N), b.std(ddof=1) / np.sqrt(N)
        se = np.hypot(sa, sb)
        # the GPU's own D row at step 10,000 is this mean (the same terms)
        g_row = r["gpu"]["dcoeff"][1, -1] / r["gpu"]["remain"][-1]
        assert abs(g_row / da - 1.0) < 1e-12
        # the reference's printed row (%10.4e) is its mean to 5 digits
        assert abs(r["ref_diff"][-1, 1] / r["ref_diff"][-1, 8] / db - 1.0) < 1e-4
        print(f"seed {seed}: D_xx GPU {da:.5e} reference {db:.5e} diff {da - db:+.3e} "
              f"= {(da - db) / se:+.2f} SE (SE {se:.3e}, {se / db:.2%})", flush=True)
        assert abs(da - db) <= K_SE * se
        d_ref.append(db)
        ses.append(sb)
    spread = float(np.std(d_ref, ddof=1))
    print(f"seed-to-seed spread of the reference's D_xx {spread:.3e} "
          f"({spread / np.mean(d_ref):.2%}); mean SE {np.mean(ses):.3e}", flush=True)
