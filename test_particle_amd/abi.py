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
# TPGPU_LIB selects another build of the library (A/B timing of variants)
LIB_PATH = os.environ.get("TPGPU_LIB") or os.path.join(HERE, "libtpgpu.so")
# the HDF5 PIC field input (include/tp_h5.h), linked to libtpgpu.so beside it
H5_LIB_PATH = os.path.join(HERE, "libtph5.so")

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

# cell-major copy of PIC grids (tp_gpu_set_grid_cells)
TP_CELLS_AUTO = 0
TP_CELLS_OFF = 1
TP_CELLS_ON = 2

# plane orders of the PIC field input (tp_gpu_load_field_grid)
TP_PLANES_X_FASTEST = 0
TP_PLANES_X_SLOWEST = 1

# microbenchmarks of the device's achievable peaks (tp_peak_measure)
TP_PEAK_HBM_COPY = 0
TP_PEAK_HBM_READ = 1
TP_PEAK_FP64_FMA = 2
TP_PEAK_FP32_FMA = 3
TP_PEAK_BF16_MFMA = 4

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
TP_PMAP_REF_COMPAT = 0
TP_PMAP_PERIODIC = 1


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


class TpBsParams(ctypes.Structure):
    _fields_ = [
        ("rtol", ctypes.c_double),
        ("atol", ctypes.c_double),
        ("qm", ctypes.c_double),
        ("t_end", ctypes.c_double),
        ("t_end_gyro", ctypes.c_double),
        ("b_ref", ctypes.c_double),
        ("h0", ctypes.c_double),
        ("kmax", ctypes.c_int),
        ("max_attempts", ctypes.c_int),
        ("block", ctypes.c_int),
        ("waves_per_cu", ctypes.c_int),
        ("use_graph", ctypes.c_int),
    ]


class TpBsStats(ctypes.Structure):
    _fields_ = [
        ("accepted", ctypes.c_longlong),
        ("rejected", ctypes.c_longlong),
        ("rhs_evals", ctypes.c_longlong),
        ("failed", ctypes.c_int),
        ("launches", ctypes.c_int),
        ("kernel_ms", ctypes.c_double),
    ]


TP_MLP_NFREQ = 128
TP_MLP_NPARAMS = 66307


class TpMlpTrainParams(ctypes.Structure):
    _fields_ = [
        ("nsamples", ctypes.c_longlong),
        ("batch", ctypes.c_int),
        ("epochs", ctypes.c_int),
        ("period", ctypes.c_double),
        ("lr", ctypes.c_double),
        ("lr_min", ctypes.c_double),
        ("beta1", ctypes.c_double),
        ("beta2", ctypes.c_double),
        ("eps", ctypes.c_double),
        ("seed", ctypes.c_ulonglong),
    ]


class TpMlpTrainStats(ctypes.Structure):
    _fields_ = [
        ("first_loss", ctypes.c_double),
        ("last_loss", ctypes.c_double),
        ("steps", ctypes.c_longlong),
        ("sample_ms", ctypes.c_double),
        ("train_ms", ctypes.c_double),
    ]


class TpPmapParams(ctypes.Structure):
    _fields_ = [
        ("face", ctypes.c_int),
        ("ncross", ctypes.c_int),
        ("mode", ctypes.c_int),
        ("chunk", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("step", ctypes.c_double),
        ("length_max", ctypes.c_double),
        ("lo", ctypes.c_double * 3),
        ("hi", ctypes.c_double * 3),
        ("lanes", ctypes.c_int),
    ]


def make_pmap_params(*, face, ncross, step, length_max, lo, hi, mode=TP_PMAP_REF_COMPAT,
                     chunk=0, precision=TP_FP64, lanes=0):
    """tp_pmap_params of a Poincare map (the reference's "Poincare map
    configuration" block: face, step, maximum length, crossing count; lo/hi
    the simulation domain)."""
    p = TpPmapParams()
    p.face, p.ncross, p.mode, p.chunk = int(face), int(ncross), int(mode), int(chunk)
    p.precision, p.lanes = int(precision), int(lanes)
    p.step, p.length_max = float(step), float(length_max)
    for d in range(3):
        p.lo[d] = float(lo[d])
        p.hi[d] = float(hi[d])
    return p


def make_mlp_train_params(*, nsamples, period, batch=16384, epochs=4, lr=2e-3, lr_min=2e-5,
                          beta1=0.9, beta2=0.999, eps=1e-8, seed=1):
    p = TpMlpTrainParams()
    p.nsamples, p.batch, p.epochs, p.period = int(nsamples), int(batch), int(epochs), period
    p.lr, p.lr_min, p.beta1, p.beta2, p.eps, p.seed = lr, lr_min, beta1, beta2, eps, seed
    return p


def make_bs_params(*, rtol=1e-8, atol=1e-10, qm=1.0, t_end=0.0, t_end_gyro=0.0, b_ref=1.0,
                   h0=0.0, kmax=8, max_attempts=0, block=0, waves_per_cu=0, use_graph=1):
    p = TpBsParams()
    p.rtol, p.atol, p.qm = rtol, atol, qm
    p.t_end, p.t_end_gyro, p.b_ref, p.h0 = t_end, t_end_gyro, b_ref, h0
    p.kmax, p.max_attempts, p.block, p.waves_per_cu, p.use_graph = (
        kmax, max_attempts, block, waves_per_cu, use_graph)
    return p


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
