"""Host-side construction of the learned field surrogate (no counterpart in
the reference; the network and its kernels are in csrc/tp_mlp.hip).

The surrogate stands in for the multi-mode turbulent field (fields.py,
``turbulence_modes``).  A coordinate network can only be periodic if the
field is, so the surrogate's field is the same mode table with every
wavevector moved to the nearest vector of the lattice 2 pi n / L of a box of
side L = 2 pi / k_min (``periodic_modes``): the smallest |k| is then one
lattice step and the field and network are both periodic in L, so particles
can wrap periodically in that box.  Amplitudes keep their length and are
re-projected perpendicular to the moved wavevector (divergence free).

Network inputs are the Fourier features [sin 2 pi (q_j . x), cos 2 pi
(q_j . x)] with q_j = n_j / L, n_j integer vectors drawn from a rounded
N(0, sigma^2) per component (random Fourier features, Tancik et al. 2020,
restricted to the period lattice).
"""
import numpy as np

from . import abi
from .fields import turbulence_modes

NFREQ = abi.TP_MLP_NFREQ
NPARAMS = abi.TP_MLP_NPARAMS
LAYERS = ((256, 128), (128, 128), (128, 128), (128, 3))


def periodic_modes(rows, period):
    """Mode table (kx ky kz ax ay az phi) with k moved to the lattice
    2 pi n / period (n integer, never zero) and a re-projected onto the
    plane perpendicular to the new k with its length kept."""
    rows = np.array(rows, dtype=np.float64, copy=True)
    step = 2.0 * np.pi / period
    n = np.rint(rows[:, :3] / step)
    zero = ~np.any(n != 0, axis=1)
    n[zero] = [0.0, 0.0, 1.0]
    k = n * step
    kh = k / np.linalg.norm(k, axis=1, keepdims=True)
    a = rows[:, 3:6]
    amp = np.linalg.norm(a, axis=1)
    ap = a - np.sum(a * kh, axis=1, keepdims=True) * kh
    norm = np.linalg.norm(ap, axis=1)
    # amplitude parallel to the moved k (rare): any perpendicular direction
    bad = norm < 1e-12 * np.maximum(amp, 1e-300)
    if np.any(bad):
        ref = np.where(np.abs(kh[bad, 2:3]) < 0.9, [[0.0, 0.0, 1.0]], [[1.0, 0.0, 0.0]])
        ap[bad] = np.cross(kh[bad], ref)
        norm[bad] = np.linalg.norm(ap[bad], axis=1)
    ap *= (amp / norm)[:, None]
    rows[:, :3] = k
    rows[:, 3:6] = ap
    return rows


def surrogate_field(nmodes=1000, seed=20261202, **kw):
    """(rows, period): the turbulence of the adaptive-stepper workload
    (fields.turbulence_modes) on the lattice of its own k_min."""
    rows = turbulence_modes(nmodes, seed=seed, **kw)
    kmin = float(np.min(np.linalg.norm(rows[:, :3], axis=1)))
    period = 2.0 * np.pi / kmin
    return periodic_modes(rows, period), period


def output_scale(rows):
    """rms of one Cartesian component of the fluctuation: the amplitude
    normalisation of the network output (sum |a|^2 / 2 over modes is the
    mean of |dB|^2)."""
    rows = np.asarray(rows, dtype=np.float64)
    return float(np.sqrt(0.5 * np.sum(rows[:, 3:6] ** 2) / 3.0))


def fourier_frequencies(period, sigma=4.0, nfreq=NFREQ, seed=0):
    """q (nfreq, 3) float32: integer lattice vectors ~ round(N(0, sigma^2)),
    zero vectors redrawn, divided by the period."""
    rng = np.random.default_rng(seed)
    n = np.rint(rng.normal(0.0, sigma, (nfreq, 3)))
    while True:
        zero = ~np.any(n != 0, axis=1)
        if not zero.any():
            break
        n[zero] = np.rint(rng.normal(0.0, sigma, (int(zero.sum()), 3)))
    return (n / period).astype(np.float32)


def fourier_frequencies_logradial(period, rmax, nfreq=NFREQ, seed=0):
    """q (nfreq, 3) float32: lattice vectors round(r d) with |r| log-uniform
    in [1, rmax] and d isotropic (zero vectors redrawn), divided by the
    period -- frequencies spread over the decades the mode table spans."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < nfreq:
        r = np.exp(rng.uniform(0.0, np.log(rmax)))
        d = rng.normal(size=3)
        n = np.rint(r * d / np.linalg.norm(d))
        if np.any(n != 0):
            out.append(n)
    return (np.array(out) / period).astype(np.float32)


def spectral_tail(rows, period, cut):
    """rms |dB| of the modes whose lattice radius |n| = |k| L / 2 pi exceeds
    cut, relative to the rms |B| (guide field 1): the part of the field a
    network whose features reach |n| <= cut cannot represent linearly."""
    rows = np.asarray(rows, dtype=np.float64)
    n = np.linalg.norm(rows[:, :3], axis=1) * period / (2.0 * np.pi)
    a2 = np.sum(rows[:, 3:6] ** 2, axis=1)
    return float(np.sqrt(0.5 * a2[n > cut].sum()) / np.sqrt(1.0 + 0.5 * a2.sum()))


def init_params(seed=0):
    """Flat float32 parameters, every weight and bias uniform in
    +-1/sqrt(fan_in) (the default of common deep-learning libraries)."""
    rng = np.random.default_rng(seed)
    out = []
    for fin, fout in LAYERS:
        b = 1.0 / np.sqrt(fin)
        out.append(rng.uniform(-b, b, fin * fout))
        out.append(rng.uniform(-b, b, fout))
    p = np.concatenate(out).astype(np.float32)
    assert p.size == NPARAMS
    return p


def split_params(p):
    """Views (W1, b1, W2, b2, W3, b3, W4, b4) of the flat vector, W[in][out]."""
    p = np.asarray(p)
    out, o = [], 0
    for fin, fout in LAYERS:
        out.append(p[o:o + fin * fout].reshape(fin, fout))
        o += fin * fout
        out.append(p[o:o + fout])
        o += fout
    return out
