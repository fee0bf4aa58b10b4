"""Microbenchmarks of the achievable peaks (include/tp_peak.h): each runs,
returns a rate that is positive and not above MI355X's published peak, and
bad arguments are refused before any launch."""
import pytest

from test_particle_amd import abi
from test_particle_amd.gpu import TPError, measure_peak
from test_particle_amd.peaks import KINDS, PUBLISHED

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(KINDS))
def test_peak_kernels_run(name):
    # the HBM sweeps on 1 GiB buffers: a buffer that fits the 256 MB
    # Infinity Cache is partly served from it on the second pass, faster
    # than HBM (8.18 TB/s was measured for a 256 MiB copy)
    nbytes = (1 << 30) if name.startswith("hbm") else (256 << 20)
    v, ms = measure_peak(KINDS[name], nbytes, 2)
    peak, _ = PUBLISHED[name]
    print(f"[peak] {name}: {v:.1f} ({ms:.3f} ms)", flush=True)
    assert ms > 0 and 0 < v <= 1.02 * peak


def test_peak_arguments():
    for what, nbytes, reps in [(5, 0, 1), (-1, 0, 1), (abi.TP_PEAK_HBM_COPY, 0, 0),
                               (abi.TP_PEAK_HBM_READ, 4096, 1)]:
        with pytest.raises(TPError):
            measure_peak(what, nbytes, reps)
    with pytest.raises(TPError):
        measure_peak(abi.TP_PEAK_FP64_FMA, 0, 1, device=99)
