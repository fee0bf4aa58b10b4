"""Seeded particle initialisation (host side).

The reference draws initial particles with drand48 seeded by time(NULL)
(init_ptl, src/particle_info.c:303-363), so its runs are not reproducible.
This module draws the same distributions from a seeded numpy Generator
(PCG64):

* positions uniform in the injection box (src/particle_info.c:336-338);
* ``isotropic``: the ``is_single_vel`` branch (src/particle_info.c:345-352):
  uz = uthe (2r - 1), phi = 2 pi r, u_perp = sqrt(uthe^2 - uz^2), i.e. an
  isotropic direction with |u| = uthe = vthe / sqrt(1 - vthe^2)
  (src/particle_info.c:317);
* ``system0``: the system-0 branch (src/particle_info.c:320-332):
  ux = uy = (i+1) uthe / n, uz = 0;
* ``loguniform``: |v| log-uniform in [vmin, vmax] with isotropic directions
  (no counterpart in the reference; used for uneven adaptive step counts).

Particles are float64 arrays of shape (n, 11) in the field order of the
reference's struct particles.
"""
import numpy as np

from .abi import NFIELD


def uthe_of(vthe):
    """gamma * v for speed vthe (src/particle_info.c:317)."""
    return vthe / np.sqrt(1.0 - vthe * vthe)


def init_particles(n, lo, hi, vthe, qm=1.0, seed=20261202, kind="isotropic",
                   vmax=None):
    rng = np.random.default_rng(seed)
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    p = np.zeros((n, NFIELD), dtype=np.float64)
    r = rng.random((n, 3))
    p[:, 0] = (hi[0] - lo[0]) * r[:, 0] + lo[0]
    p[:, 1] = (hi[1] - lo[1]) * r[:, 1] + lo[1]
    p[:, 2] = (hi[2] - lo[2]) * r[:, 2] + lo[2]
    p[:, 3] = qm
    if kind == "system0":
        u = uthe_of(vthe)
        k = np.arange(1, n + 1, dtype=np.float64)
        p[:, 4] = k * u / n
        p[:, 5] = k * u / n
        return p
    if kind == "isotropic":
        u = np.full(n, uthe_of(vthe))
    elif kind == "loguniform":
        v = np.exp(rng.uniform(np.log(vthe), np.log(vmax), n))
        u = uthe_of(v)
    else:
        raise ValueError(kind)
    r2 = rng.random((n, 2))
    uz = u * (r2[:, 0] * 2 - 1)
    phi = r2[:, 1] * np.pi * 2
    vxy = np.sqrt(np.maximum(u * u - uz * uz, 0.0))
    p[:, 4] = vxy * np.cos(phi)
    p[:, 5] = vxy * np.sin(phi)
    p[:, 6] = uz
    return p
