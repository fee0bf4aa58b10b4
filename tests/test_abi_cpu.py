"""CPU-side checks of the boundary: libtpgpu.so and libtph5.so load and
export every function their headers in include/ declare; the Python mirror's structures match the
C layout."""
import ctypes
import os
import re

import pytest

from test_particle_amd import abi

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
HEADERS = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
           if f.endswith(".h")]


# include/tp_h5.h is the HDF5 input library's; every other header is libtpgpu.so's
H5_HEADER = os.path.join(ROOT, "include", "tp_h5.h")
H5_LIB = os.path.join(os.path.dirname(abi.LIB_PATH), "libtph5.so")


def declared_functions(headers=None):
    names = set()
    for h in headers or [h for h in HEADERS if h != H5_HEADER]:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(?:int|void|const char\s*\*)\s*\**\s*(tp_\w+)\s*\(", src):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("tp_gpu_init", "tp_gpu_track_fixed", "tp_gpu_set_field_grid",
                 "tp_gpu_run_fixed", "tp_gpu_finalize"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(abi.LIB_PATH), reason="libtpgpu.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


@pytest.mark.skipif(not os.path.exists(H5_LIB), reason="libtph5.so not built")
def test_h5_library_exports_every_declared_symbol():
    names = declared_functions([H5_HEADER])
    assert "tp_h5_load_pic" in names and "tp_h5_read_pic" in names
    lib = ctypes.CDLL(H5_LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_particle_struct_is_88_bytes():
    assert abi.NFIELD * 8 == 88


def test_track_params_layout():
    # 10 ints followed by 10 doubles, naturally aligned
    assert ctypes.sizeof(abi.TpTrackParams) == 10 * 4 + 10 * 8
    assert abi.TpTrackParams.dt.offset == 40


def test_make_params_follows_reference_derivations():
    p = abi.make_params(dt=0.1, estep=10000, lo=(0, 0, 0), hi=(1, 1, 1), nbins=102,
                        nt_out=100, emin=1e-3, emax=10.0)
    assert p.einterval_t == 100                       # estep / nt_out
    assert abs(p.logde - 4.0 / 100) < 1e-15           # (log emax - log emin)/(nbins-2)
    assert abi.n_dcoeff_rows(p) == 101                # estep/100 + 1


@pytest.mark.skipif(not os.path.exists(abi.LIB_PATH), reason="libtpgpu.so not built")
def test_strerror_without_gpu():
    lib = ctypes.CDLL(abi.LIB_PATH)
    lib.tp_gpu_strerror.restype = ctypes.c_char_p
    assert lib.tp_gpu_strerror(-1) == b"invalid argument"
    assert lib.tp_gpu_strerror(0) == b"ok"
