"""The failure the library's TP_FLAG_CHECK_FINITE reports, shown on the CPU.

The reference's Wirz push divides by |B| (iBtot = 1/|B|,
src/tracking.c:782), so a particle where B = 0 gets infinite and then NaN
state, which the reference carries on silently.  The oracle restates that
push (it is pinned bitwise to the reference's golden vectors), so here it
shows the condition the GPU path turns into TP_E_NONFINITE; the Boris push
has no such division and stays finite."""
import numpy as np

from test_particle_amd import abi
from test_particle_amd.particles import init_particles

TWO_PI = 2.0 * np.pi


def test_wirz_in_zero_field_is_not_finite(oracle_tp):
    ptl = init_particles(8, (0, 0, 0), (TWO_PI,) * 3, 0.7, seed=5)
    zero = oracle_tp.Field.abc(1.0, 1.0, 1.0, B0=0.0)
    out = oracle_tp.steps(zero, ptl, 3, 0.05, pusher=abi.TP_PUSH_WIRZ)
    assert not np.isfinite(out[:, [0, 1, 2, 4, 5, 6]]).all(axis=1).any()


def test_boris_in_zero_field_stays_finite(oracle_tp):
    ptl = init_particles(8, (0, 0, 0), (TWO_PI,) * 3, 0.7, seed=5)
    zero = oracle_tp.Field.abc(1.0, 1.0, 1.0, B0=0.0)
    out = oracle_tp.steps(zero, ptl, 3, 0.05, pusher=abi.TP_PUSH_BORIS)
    assert np.isfinite(out).all()
    # no field: straight lines at constant momentum
    assert np.array_equal(out[:, 4:7], ptl[:, 4:7])
