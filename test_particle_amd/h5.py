"""PIC field input from the reference's HDF5 files (include/tp_h5.h over
libtph5.so).

The reference's system-4 run reads "<path>/T.<tf>/fields_<tf>.h5", group
"Timestep_<tf>" (src/domain.c:384-386), datasets ex ey ez cbx cby cbz stored
x slowest (read_pic_efields_float_h5, src/electric_field.c:447-507;
read_pic_bfields_float_h5, src/magnetic_field.c:632-692).  ``load_pic``
streams them onto the GPU; ``read_pic`` returns the host arrays the
reference's loaders build.
"""
import ctypes
import os

import numpy as np

from . import abi
from .abi import fptr
from .gpu import _check, _vec3, lib as _gpu_lib

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(abi.H5_LIB_PATH):
            raise RuntimeError(
                f"{abi.H5_LIB_PATH} is not built; run `python -m test_particle_amd.build`")
        _gpu_lib()   # the same libtpgpu.so instance the contexts come from
        L = ctypes.CDLL(abi.H5_LIB_PATH)
        P = ctypes.POINTER
        d, i, vp, cs, f = (ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p,
                           ctypes.c_float)
        L.tp_h5_read_pic_planes.argtypes = [cs, cs, i, i, i, P(f)]
        L.tp_h5_read_pic.argtypes = [cs, cs, i, i, i, d, d, P(f), P(f)]
        L.tp_h5_load_pic.argtypes = [vp, cs, cs, i, i, i, P(d), P(d), d, d, i]
        L.tp_h5_write_pic_planes.argtypes = [cs, cs, P(ctypes.c_longlong), P(f)]
        L.tp_h5_write_particles_fields.argtypes = [cs, i, P(d), P(d)]
        L.tp_h5_save_particles_fields.argtypes = [vp, cs, i, P(d)]
        L.tp_h5_read_particles_fields.argtypes = [cs, i, P(i), P(d), P(d)]
        _lib = L
    return _lib


def pic_file(path, tf):
    """(file name, group name) of time frame tf, as src/domain.c:384-386
    forms them."""
    return f"{path}/T.{tf}/fields_{tf}.h5", f"Timestep_{tf}"


def read_pic_planes(fname, gname, shape):
    """The six datasets' leading (nx, ny, nz) blocks: float32 (6, nx, ny, nz),
    unscaled."""
    nx, ny, nz = (int(v) for v in shape)
    out = np.zeros((6, nx, ny, nz), np.float32)
    _check(lib().tp_h5_read_pic_planes(fname.encode(), gname.encode(), nx, ny, nz, fptr(out)),
           "h5_read_pic_planes")
    return out


def read_pic(fname, gname, shape, e0, b0):
    """(e, b) as the reference's loaders leave efc_float and magc_float:
    float32 (nz, ny, nx, 3), scaled by e0 and b0."""
    nx, ny, nz = (int(v) for v in shape)
    e = np.zeros((nz, ny, nx, 3), np.float32)
    b = np.zeros((nz, ny, nx, 3), np.float32)
    _check(lib().tp_h5_read_pic(fname.encode(), gname.encode(), nx, ny, nz, float(e0),
                                float(b0), fptr(e), fptr(b)), "h5_read_pic")
    return e, b


def load_pic(tracker, fname, gname, shape, lo, hi, e0, b0, mode=abi.TP_GATHER_REF_COMPAT):
    """Streams the file onto the tracker's GPU (tp_h5_load_pic)."""
    nx, ny, nz = (int(v) for v in shape)
    _check(lib().tp_h5_load_pic(tracker._ctx, fname.encode(), gname.encode(), nx, ny, nz,
                                _vec3(lo), _vec3(hi), float(e0), float(b0), mode),
           "h5_load_pic")


def write_pic_planes(fname, gname, planes):
    """Writes planes (6, x, y, z) as the datasets ex ey ez cbx cby cbz of a new
    file (tp_h5_write_pic_planes)."""
    p = np.ascontiguousarray(planes, dtype=np.float32)
    if p.ndim != 4 or p.shape[0] != 6:
        raise ValueError("planes must have shape (6, nx, ny, nz)")
    dims = np.array(p.shape[1:], dtype=np.int64)
    _check(lib().tp_h5_write_pic_planes(fname.encode(), gname.encode(),
                                        dims.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                        fptr(p)), "h5_write_pic_planes")


def write_particles_fields(fname, ptl, emf):
    """data/particle_diagnostics.h5 as trajectory_diagnostics writes it
    (tp_h5_write_particles_fields): the trajectory points ptl and the field
    emf[i] = (Bx, By, Bz, Ex, Ey, Ez) at each."""
    ptl = abi.as_particles(ptl)
    emf = np.ascontiguousarray(emf, dtype=np.float64).reshape(-1, 6)
    if emf.shape[0] != ptl.shape[0]:
        raise ValueError("one field row per point")
    _check(lib().tp_h5_write_particles_fields(os.fsencode(fname), ptl.shape[0], abi.dptr(ptl),
                                              abi.dptr(emf)), "h5_write_particles_fields")


def save_particles_fields(tracker, fname, ptl):
    """The same with the field evaluated on the tracker's GPU at each point
    (tp_h5_save_particles_fields)."""
    ptl = abi.as_particles(ptl)
    _check(lib().tp_h5_save_particles_fields(tracker._ctx, os.fsencode(fname), ptl.shape[0],
                                             abi.dptr(ptl)), "h5_save_particles_fields")


def read_particles_fields(fname, cap=1 << 20):
    """(points [n, 8] = x y z qm ux uy uz t, fields [n, 6]) of such a file."""
    p8 = np.zeros((cap, 8))
    emf = np.zeros((cap, 6))
    n = ctypes.c_int(0)
    _check(lib().tp_h5_read_particles_fields(os.fsencode(fname), cap, ctypes.byref(n),
                                             abi.dptr(p8), abi.dptr(emf)),
           "h5_read_particles_fields")
    return p8[:n.value].copy(), emf[:n.value].copy()
