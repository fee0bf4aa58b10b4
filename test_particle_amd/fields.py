"""Host-side builders of the field sources (inputs to the tracking path).

* ``mode_table``: the multi-mode turbulence of the adaptive and surrogate
  workloads (no counterpart in the reference): B = b0z z-hat +
  sum_n a_n cos(k_n . x + phi_n) with |k_n| log-spaced over [kmin, kmax],
  k-hat_n isotropic, a_n perpendicular to k_n (divergence free), random
  polarisation angle and phase, |a_n| proportional to |k_n|^index, scaled so
  that the spatial rms of |dB| equals ``rms`` (rms^2 = sum |a_n|^2 / 2).
* ``pic_turbulence_grid``: the gridded PIC-field workload.  A Gaussian random
  field with an isotropic k^-5/3 omnidirectional spectrum (3D amplitude
  ~ k^-11/6), random phases and a solenoidal projection, after the recipe of
  the reference's generator (data/turbulence.c:228-257, 310-321), built with
  FFTs on a periodic n^3 lattice (spacing 1) and stored on (n+1)^3 nodes with
  the first plane repeated last, because the reference's grid is
  endpoint-inclusive (dx = L/(nx-1), src/domain.c:548) and wraps particles at
  xmax (src/tracking.c:1186-1210).  B = b0 z-hat + dB (rms |dB| = db_rms);
  E = -u x B from an independent velocity field of the same spectrum with
  rms |u| = u_rms.
* ``abc_gyro_dt``: dt as a fraction of the gyroperiod T_g = 2 pi gamma0 /
  (|q/m| B_ref).
"""
import numpy as np


def mode_table(nmodes, kmin, kmax, rms, seed=0, index=-5.0 / 3.0):
    rng = np.random.default_rng(seed)
    if nmodes == 1:
        k = np.array([kmin], dtype=np.float64)
    else:
        k = kmin * (kmax / kmin) ** (np.arange(nmodes) / (nmodes - 1.0))
    ct = rng.uniform(-1.0, 1.0, nmodes)
    ph = rng.uniform(0.0, 2.0 * np.pi, nmodes)
    st = np.sqrt(1.0 - ct * ct)
    kh = np.column_stack([st * np.cos(ph), st * np.sin(ph), ct])
    # orthonormal pair perpendicular to k-hat
    ref = np.where(np.abs(kh[:, 2:3]) < 0.9, [[0.0, 0.0, 1.0]], [[1.0, 0.0, 0.0]])
    e1 = np.cross(kh, ref)
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = np.cross(kh, e1)
    psi = rng.uniform(0.0, 2.0 * np.pi, nmodes)
    ehat = np.cos(psi)[:, None] * e1 + np.sin(psi)[:, None] * e2
    amp = k ** index
    amp *= rms / np.sqrt(0.5 * np.sum(amp * amp))
    phase = rng.uniform(0.0, 2.0 * np.pi, nmodes)
    rows = np.column_stack([kh * k[:, None], ehat * amp[:, None], phase])
    return np.ascontiguousarray(rows, dtype=np.float64)


def turbulence_modes(nmodes=1000, decades=3.0, kmin_rg=0.1, vmax=0.9, rms=0.5, b0=1.0,
                     seed=20261202, index=-5.0 / 3.0):
    """The adaptive-stepper workload's field: nmodes modes with k log-spaced
    over `decades` decades from kmin = kmin_rg / r_g(vmax), r_g = gamma v /
    (|q/m| B0) with q/m = 1; rms |dB| = rms over the guide field b0."""
    u = vmax / np.sqrt(1.0 - vmax * vmax)
    kmin = kmin_rg / (u / b0)
    return mode_table(nmodes, kmin, kmin * 10.0 ** decades, rms, seed=seed, index=index)


def eval_modes_numpy(rows, pts, b0z):
    """Direct FP64 sum of the mode table at pts (n, 3): the brute-force
    check of the in-kernel evaluator."""
    ph = pts @ rows[:, :3].T + rows[:, 6]
    c = np.cos(ph)
    b = c @ rows[:, 3:6]
    b[:, 2] += b0z
    return b


def _solenoidal_gaussian(n, rng, index_amp, workers=-1):
    """Three real components of a divergence-free Gaussian random field on
    an n^3 periodic lattice, amplitude ~ |k|^index_amp."""
    import scipy.fft as sfft
    k1 = 2.0 * np.pi * np.fft.fftfreq(n)
    kr = 2.0 * np.pi * np.fft.rfftfreq(n)
    kz, ky, kx = np.meshgrid(k1, k1, kr, indexing="ij")
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where(kk > 0, kk ** index_amp, 0.0)
        inv = np.where(kk > 0, 1.0 / kk, 0.0)
    khat = (kx * inv, ky * inv, kz * inv)
    comps = []
    for _ in range(3):
        w = rng.standard_normal((n, n, n))
        comps.append(sfft.rfftn(w, workers=workers) * amp)
    dot = comps[0] * khat[0] + comps[1] * khat[1] + comps[2] * khat[2]
    out = []
    for a in range(3):
        f = comps[a] - dot * khat[a]
        out.append(sfft.irfftn(f, s=(n, n, n), workers=workers))
    return np.stack(out, axis=-1)      # (z, y, x, 3)


def pic_turbulence_grid(n=256, seed=20261202, db_rms=1.0, u_rms=0.05, b0=1.0,
                        index=-5.0 / 3.0):
    """Returns (efield, bfield, lo, hi): float32 arrays of shape
    (n+1, n+1, n+1, 3) over [0, n]^3."""
    rng = np.random.default_rng(seed)
    index_amp = (index - 2.0) / 2.0       # E(k) ~ k^index -> |f(k)| ~ k^((index-2)/2)
    db = _solenoidal_gaussian(n, rng, index_amp)
    db *= db_rms / np.sqrt(np.mean(np.sum(db * db, axis=-1)))
    u = _solenoidal_gaussian(n, rng, index_amp)
    u *= u_rms / np.sqrt(np.mean(np.sum(u * u, axis=-1)))
    b = db
    b[..., 2] += b0
    e = -np.cross(u, b)
    del u

    def ext(a):
        a = np.concatenate([a, a[:1]], axis=0)
        a = np.concatenate([a, a[:, :1]], axis=1)
        a = np.concatenate([a, a[:, :, :1]], axis=2)
        return np.ascontiguousarray(a, dtype=np.float32)

    return ext(e), ext(b), (0.0, 0.0, 0.0), (float(n), float(n), float(n))


def gyroperiod(vthe, b_ref, qm=1.0):
    """T_g = 2 pi gamma0 / (|q/m| B_ref)."""
    gamma0 = 1.0 / np.sqrt(1.0 - vthe * vthe)
    return 2.0 * np.pi * gamma0 / (abs(qm) * b_ref)


def abc_b_ref(A=1.0, B=1.0, C=1.0, B0=1.0):
    """rms |B| of the ABC field: B0 sqrt(A^2 + B^2 + C^2)."""
    return B0 * np.sqrt(A * A + B * B + C * C)
