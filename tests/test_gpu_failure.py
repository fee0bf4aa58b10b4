"""Error paths of the C ABI on the GPU: non-finite state reported by
TP_FLAG_CHECK_FINITE (B = 0 makes the reference's Wirz push divide by zero,
src/tracking.c:782), in the one-call path and the resident-particle path,
with the outputs still written; bad arguments refused before any launch."""
import numpy as np
import pytest

from test_particle_amd import abi
from test_particle_amd.particles import init_particles

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * np.pi
BOX = ((0.0, 0.0, 0.0), (TWO_PI, TWO_PI, TWO_PI))


@pytest.fixture(scope="module")
def gpu():
    from test_particle_amd.gpu import Tracker
    t = Tracker(0)
    yield t
    t.close()


def _params(flags, pusher=abi.TP_PUSH_WIRZ):
    return abi.make_params(dt=0.05, estep=200, lo=BOX[0], hi=BOX[1], pusher=pusher,
                           nbins=102, nt_out=2, emin=1e-3, emax=10.0, flags=flags)


def test_zero_field_wirz_reports_nonfinite(gpu):
    from test_particle_amd.gpu import TPError
    ptl = init_particles(300, BOX[0], BOX[1], 0.7, seed=3)
    gpu.set_field_abc(1.0, 1.0, 1.0, B0=0.0)
    print("[failure] track_fixed, B = 0, Wirz, CHECK_FINITE", flush=True)
    with pytest.raises(TPError) as e:
        gpu.track_fixed(_params(abi.TP_FLAG_CHECK_FINITE), ptl)
    assert e.value.code == abi.TP_E_NONFINITE
    print("[failure] resident run, B = 0, Wirz, CHECK_FINITE", flush=True)
    gpu.load(ptl)
    with pytest.raises(TPError) as e:
        gpu.run(_params(abi.TP_FLAG_CHECK_FINITE))
    assert e.value.code == abi.TP_E_NONFINITE
    out, ns = gpu.store()
    assert not np.isfinite(out[:, [0, 1, 2, 4, 5, 6]]).all()
    assert (ns == 201).all()
    # without the flag the run completes and carries the NaNs, as the reference
    gpu.load(ptl)
    gpu.run(_params(0))
    out, _ = gpu.store()
    assert not np.isfinite(out).all()


def test_zero_field_boris_is_finite(gpu):
    ptl = init_particles(300, BOX[0], BOX[1], 0.7, seed=3)
    gpu.set_field_abc(1.0, 1.0, 1.0, B0=0.0)
    r = gpu.track_fixed(_params(abi.TP_FLAG_CHECK_FINITE, abi.TP_PUSH_BORIS), ptl)
    assert np.isfinite(r["ptl"]).all()


def test_bad_arguments_are_refused(gpu):
    from test_particle_amd.gpu import TPError
    gpu.set_field_abc(1.0, 1.0, 1.0)
    ptl = init_particles(16, BOX[0], BOX[1], 0.7, seed=4)
    gpu.load(ptl)
    for bad in (dict(dcoeff_interval=0), dict(nt_out=0, einterval_t=1), dict(einterval_t=0),
                dict(dt=0.0), dict(bc_flag=3)):
        kw = dict(dt=0.05, estep=10, lo=BOX[0], hi=BOX[1], nt_out=1)
        kw.update(bad)
        with pytest.raises(TPError) as e:
            gpu.run(abi.make_params(**kw))
        assert e.value.code == abi.TP_E_ARG
