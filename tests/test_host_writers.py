"""The host-side writers reproduce the reference's output files: the oracle
(libm mode, bitwise the reference's loop) feeds them the raw arrays of the
reference binary's own particle_tracking_hybrid run, and the text must be
identical to data/espectrum.dat and data/diff_coeffs.dat it wrote."""
import os

import numpy as np
import pytest

from conftest import golden
from test_particle_amd import abi

pytestmark = pytest.mark.skipif(not os.path.exists(abi.LIB_PATH),
                                reason="libtpgpu.so not built")


def test_reference_output_files(oracle_libm, tmp_path):
    from test_particle_amd import gpu
    g = golden("hybrid_files_abc1")
    dt = float(g["dt"])
    two_pi = 2 * np.pi
    estep = int((300 * dt + 1e-9) / dt)
    prm = abi.make_params(dt=dt, estep=estep, lo=(0, 0, 0), hi=(two_pi,) * 3, nbins=102,
                          nt_out=3, emin=1e-3, emax=10.0)
    r = oracle_libm.track_fixed(oracle_libm.Field.abc(1.0, 1.0, 1.0), prm, g["ptl"],
                                nthreads=1)
    gpu.write_outputs(str(tmp_path), g["ptl"], r, 102, 1e-3, 10.0)
    assert (tmp_path / "espectrum.dat").read_text() == str(g["espectrum_dat"])
    assert (tmp_path / "diff_coeffs.dat").read_text() == str(g["diff_coeffs_dat"])


def test_spectrum_bins_formulas():
    from test_particle_amd import gpu
    eb, ei = gpu.spectrum_bins(12, 1e-2, 1e3)
    logde = 5.0 / 10
    assert eb[0] == 0.5e-2 and ei[0] == 1e-2
    assert np.isclose(eb[3], 10 ** (-2 + logde * 2.5))
    assert np.isclose(ei[3], 1e-2 * (10 ** (logde * 3) - 10 ** (logde * 2)))


@pytest.mark.parametrize("two", [False, True])
@pytest.mark.parametrize("ct", [0, 2, 3])
def test_read_pic_gda_matches_reference_loader(tmp_path, two, ct):
    """The .gda field input (tp_host_read_pic_gda) loads exactly what the
    reference's read_efields/read_bfields_binary load: the same slices
    (previous ct - 1 and next ct + 1, clamped at the ends), the same
    double-then-float scaling, the same node order."""
    from conftest import golden
    from test_particle_amd.gpu import read_pic_gda
    g = golden("gda_slices")
    nx, ny, nz, nt = int(g["nx"]), int(g["ny"]), int(g["nz"]), int(g["nt"])
    for k in g.files:
        if k[:2] in ("ex", "ey", "ez", "bx", "by", "bz"):
            g[k].tofile(os.path.join(str(tmp_path), k + ".gda"))
    ref = g[f"mt{int(two)}_ct{ct}"].reshape(-1, nz, ny, nx, 3)
    out = read_pic_gda(str(tmp_path), ct, nx, ny, nz, float(g["e0"]), float(g["b0"]), nt=nt,
                       two_slices=two)
    assert len(out) == ref.shape[0]
    for a, b in zip(out, ref):
        assert np.array_equal(a, b)
