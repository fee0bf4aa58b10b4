"""PIC HDF5 field input on the CPU (include/tp_h5.h, libtph5.so).

The golden fixture pic_h5.npz holds a file's six datasets (9 x 6 x 8 floats,
x slowest) and what the reference's own loaders (read_pic_efields_h5 and
read_pic_bfields_h5, src/electric_field.c:447-507, src/magnetic_field.c:
632-692, run by oracle/ref_harness.c) left in efc_float and magc_float for
the leading 7 x 5 x 6 block.  tp_h5_read_pic must load the same bits.
"""
import os

import numpy as np
import pytest

from conftest import golden
from test_particle_amd import abi

pytestmark = pytest.mark.skipif(not os.path.exists(abi.H5_LIB_PATH),
                                reason="libtph5.so not built")


def _fixture_file(tmp_path):
    from test_particle_amd import h5
    g = golden("pic_h5")
    fname = str(tmp_path / "fields_3.h5")
    h5.write_pic_planes(fname, str(g["gname"]), g["planes"])
    return g, fname


def test_golden_is_transpose_and_double_scaling():
    """What the reference's loaders did, in closed form: node (i, j, k) of
    component c is (float)(dataset_c[i, j, k] * scale) in double, stored x
    fastest (an exchanged index or a float product fails this)."""
    g = golden("pic_h5")
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    blk = g["planes"][:, :nx, :ny, :nz].astype(np.float64)
    e = (blk[:3] * float(g["e0"])).astype(np.float32).transpose(3, 2, 1, 0)
    b = (blk[3:] * float(g["b0"])).astype(np.float32).transpose(3, 2, 1, 0)
    assert np.array_equal(e, g["e"]) and np.array_equal(b, g["b"])
    # the float product would differ somewhere on this data
    ef = (g["planes"][:3, :nx, :ny, :nz] * np.float32(g["e0"])).transpose(3, 2, 1, 0)
    assert not np.array_equal(ef, g["e"])


def test_read_pic_matches_reference_loaders(tmp_path):
    from test_particle_amd import h5
    g, fname = _fixture_file(tmp_path)
    shape = (int(g["nx"]), int(g["ny"]), int(g["nz"]))
    e, b = h5.read_pic(fname, str(g["gname"]), shape, float(g["e0"]), float(g["b0"]))
    assert np.array_equal(e, g["e"])
    assert np.array_equal(b, g["b"])


def test_read_planes_is_leading_block(tmp_path):
    from test_particle_amd import h5
    g, fname = _fixture_file(tmp_path)
    for shape in [(7, 5, 6), (9, 6, 8), (1, 1, 1), (3, 6, 2)]:
        p = h5.read_pic_planes(fname, str(g["gname"]), shape)
        assert np.array_equal(p, g["planes"][:, :shape[0], :shape[1], :shape[2]])


def test_errors(tmp_path):
    from test_particle_amd import h5
    from test_particle_amd.gpu import TPError
    g, fname = _fixture_file(tmp_path)
    gname = str(g["gname"])
    with pytest.raises(TPError):
        h5.read_pic(str(tmp_path / "missing.h5"), gname, (7, 5, 6), 1.0, 1.0)
    with pytest.raises(TPError):
        h5.read_pic(fname, "Timestep_4", (7, 5, 6), 1.0, 1.0)
    with pytest.raises(TPError):                       # block larger than the datasets
        h5.read_pic(fname, gname, (10, 5, 6), 1.0, 1.0)
    with pytest.raises(TPError):
        h5.read_pic_planes(fname, gname, (7, 0, 6))
    import ctypes
    L = h5.lib()                                       # no output buffer
    assert L.tp_h5_read_pic_planes(fname.encode(), gname.encode(), 2, 2, 2,
                                   ctypes.cast(0, ctypes.POINTER(ctypes.c_float))) == abi.TP_E_ARG


def test_pic_file_names():
    from test_particle_amd import h5
    assert h5.pic_file("/d/fields", 12) == ("/d/fields/T.12/fields_12.h5", "Timestep_12")
