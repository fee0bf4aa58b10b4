"""The trajectory output file on the CPU (tp_h5_write_particles_fields,
tp_h5_read_particles_fields in libtph5.so): data/particle_diagnostics.h5 as
trajectory_diagnostics writes it through save_particles_fields
(src/diagnostics.c:290-390).  The reference's writer opens the file through
MPI-IO, which the serial HDF5 here lacks, so the layout is pinned by the
reference's type definitions -- group /particles_fields, datasets
"particles" with members x y z qm ux uy uz t (src/particle_info.c:452-488)
and "fields" with Bx By Bz Ex Ey Ez (src/emfields.c:227-257), big-endian
IEEE doubles at offsets 8*k -- checked in h5dump's own description of the
file, and the values by reading them back."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from test_particle_amd import abi

pytestmark = pytest.mark.skipif(not os.path.exists(abi.H5_LIB_PATH),
                                reason="libtph5.so not built")

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                    if os.path.exists("/opt/conda/bin/h5dump") else None)


def _data(n, seed):
    rng = np.random.default_rng(seed)
    ptl = rng.normal(0, 3, (n, abi.NFIELD))
    emf = rng.normal(0, 1, (n, 6))
    return ptl, emf


def test_round_trip(tmp_path):
    from test_particle_amd import h5
    ptl, emf = _data(257, 61)
    f = str(tmp_path / "particle_diagnostics.h5")
    h5.write_particles_fields(f, ptl, emf)
    p8, e = h5.read_particles_fields(f)
    assert np.array_equal(p8, ptl[:, :8])       # the shifts are not written
    assert np.array_equal(e, emf)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not available")
def test_layout_is_the_reference_compound_types(tmp_path):
    from test_particle_amd import h5
    ptl, emf = _data(5, 62)
    f = str(tmp_path / "particle_diagnostics.h5")
    h5.write_particles_fields(f, ptl, emf)
    head = subprocess.run([H5DUMP, "-H", f], capture_output=True, text=True, timeout=60,
                          check=True).stdout
    assert 'GROUP "particles_fields"' in head       # under the root group
    for dset, names in (("particles", ["x", "y", "z", "qm", "ux", "uy", "uz", "t"]),
                        ("fields", ["Bx", "By", "Bz", "Ex", "Ey", "Ez"])):
        m = re.search(r'DATASET "%s" \{\s*DATATYPE\s+H5T_COMPOUND \{(.*?)\}\s*DATASPACE\s+'
                      r'SIMPLE \{ \( (\d+) \) / \( (\d+) \) \}' % dset, head, re.S)
        assert m, head
        members = re.findall(r'(H5T_\w+)\s+"(\w+)";', m.group(1))
        assert members == [("H5T_IEEE_F64BE", k) for k in names], members
        assert int(m.group(2)) == 5 and int(m.group(3)) == 5
    # the values, as h5dump prints them by member name, are the doubles written
    for dset, want in (("fields", emf), ("particles", ptl[:, :8])):
        vals = subprocess.run([H5DUMP, "-d", f"/particles_fields/{dset}", "-m", "%.17g", f],
                              capture_output=True, text=True, timeout=60, check=True).stdout
        data = vals[vals.index("DATA {"):]
        got = [float(v) for v in re.findall(r"[-+]?\d+\.\d+(?:e[-+]?\d+)?|[-+]?\d+e[-+]?\d+",
                                            data)]
        assert np.array_equal(np.array(got), want.ravel()), dset


def test_empty_and_bad_files(tmp_path):
    from test_particle_amd import h5
    from test_particle_amd.gpu import TPError
    f = str(tmp_path / "empty.h5")
    h5.write_particles_fields(f, np.zeros((0, abi.NFIELD)), np.zeros((0, 6)))
    p8, e = h5.read_particles_fields(f)
    assert p8.shape == (0, 8) and e.shape == (0, 6)
    # a PIC field file has no /particles_fields
    g = str(tmp_path / "fields.h5")
    h5.write_pic_planes(g, "Timestep_0", np.zeros((6, 2, 2, 2), dtype=np.float32))
    with pytest.raises(TPError):
        h5.read_particles_fields(g)
    with pytest.raises(TPError):
        h5.read_particles_fields(str(tmp_path / "missing.h5"))
    ptl, emf = _data(9, 63)
    h5.write_particles_fields(f, ptl, emf)
    with pytest.raises(TPError):                 # more entries than the capacity
        h5.read_particles_fields(f, cap=8)
    with pytest.raises(ValueError):
        h5.write_particles_fields(f, ptl, emf[:3])
