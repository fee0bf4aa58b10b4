"""ctypes mirror of include/tp_gpu.h (the C ABI of libtpgpu.so).

The structures here are byte-compatible with the C declarations; particle
arrays are handled as contiguous float64 arrays of shape (n, 11) in the
field order of the reference's ``struct particles``
(include/particle_info.h:6-12).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtpgpu.so")

PARTICLE_FIELDS = ("x", "y", "z", "qm", "ux", "uy", "uz", "t",
                   "xshift", "yshift", "zshift")
NFIELD = len(PARTICLE_FIELDS)

TP_OK = 0
TP_E_ARG = -1
TP_E_NOMEM = -2
TP_E_HIP = -3
TP_E_NONFINITE = -4
TP_E_STATE = -5
TP_E_NODEV = -6

TP_PUSH_WIRZ = 0
TP_PUSH_BORIS = 1
TP_FP64 = 0
TP_FP32_MIXED = 1
TP_BC_PERIODIC = 0
TP_BC_OPEN = 1
TP_BC_NONE = 2
TP_GATHER_REF_COMPAT = 0
TP_GATHER_TEXTBOOK = 1
TP_FLAG_FIX_STALE_DCOEFF = 1
TP_FLAG_CHECK_FINITE = 2


class TpTrackParams(ctypes.Structure):
    _fields_ = [
        ("pusher", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("bc_flag", ctypes.c_int),
        ("nbins", ctypes.c_int),
        ("nt_out", ctypes.c_int),
        ("sstep", ctypes.c_int),
        ("estep", ctypes.c_int),
        ("einterval_t", ctypes.c_int),
        ("dcoeff_interval", ctypes.c_int),
        ("flags", ctypes.c_int),
        ("dt", ctypes.c_double),
        ("qm", ctypes.c_double),
        ("xmin", ctypes.c_double),
        ("xmax", ctypes.c_double),
        ("ymin", ctypes.c_double),
        ("ymax", ctypes.c_double),
        ("zmin", ctypes.c_double),
        ("zmax", ctypes.c_double),
        ("logemin", ctypes.c_double),
        ("logde", ctypes.c_double),
    ]


def make_params(*, dt, estep, lo, hi, qm=1.0, pusher=TP_PUSH_WIRZ,
                precision=TP_FP64, bc_flag=TP_BC_PERIODIC, nbins=102,
                nt_out=1, emin=1e-3, emax=1e1, sstep=1, einterval_t=None,
                dcoeff_interval=100, flags=0):
    """Parameters as the reference derives them: einterval_t = estep/nt_out
    (src/tracking.c:104), logemin/logde from emin/emax/nbins
    (src/diagnostics.c:629-631)."""
    p = TpTrackParams()
    p.pusher = pusher
    p.precision = precision
    p.bc_flag = bc_flag
    p.nbins = nbins
    p.nt_out = nt_out
    p.sstep = sstep
    p.estep = int(estep)
    p.einterval_t = int(estep) // nt_out if einterval_t is None else einterval_t
    p.dcoeff_interval = dcoeff_interval
    p.flags = flags
    p.dt = float(dt)
    p.qm = float(qm)
    p.xmin, p.ymin, p.zmin = (float(v) for v in lo)
    p.xmax, p.ymax, p.zmax = (float(v) for v in hi)
    logemin = np.log10(emin)
    p.logemin = float(logemin)
    p.logde = float((np.log10(emax) - logemin) / (nbins - 2))
    return p


def n_dcoeff_rows(p):
    """Rows of the diffusion-coefficient arrays: estep/100 + 1
    (src/tracking.c:109)."""
    return p.estep // p.dcoeff_interval + 1


def as_particles(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] != NFIELD:
        raise ValueError("particles must be float64 of shape (n, 11)")
    return a


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) if a is not None else None
