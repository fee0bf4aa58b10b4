"""ctypes wrapper of the CPU oracle libtpref.so -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
as the checker; never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

from test_particle_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libtpref.so")
REF_DIR = os.path.join(HERE, "_ref")

FIELD_TEST, FIELD_WLCS, FIELD_ABC, FIELD_GRID, FIELD_MODES = 0, 1, 2, 4, 5
FIELD_MHD, FIELD_GRIDD = 3, 6


class TprefField(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("abc", ctypes.c_double * 5),
        ("vnx", ctypes.c_int), ("vny", ctypes.c_int), ("vnz", ctypes.c_int),
        ("vlo", ctypes.c_double * 3), ("vhi", ctypes.c_double * 3),
        ("vgrid", ctypes.POINTER(ctypes.c_double)),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3),
        ("egrid", ctypes.POINTER(ctypes.c_float)),
        ("bgrid", ctypes.POINTER(ctypes.c_float)),
        ("gather_mode", ctypes.c_int),
        ("nmodes", ctypes.c_int),
        ("modes", ctypes.POINTER(ctypes.c_double)),
        ("b0z", ctypes.c_double),
        ("nwlcs", ctypes.c_int),
        ("wlcs_rows", ctypes.POINTER(ctypes.c_double)),
        ("ibessel", ctypes.c_int),
        ("omega", ctypes.c_double),
        ("norm_i", ctypes.c_double),
        ("wlcs_mode", ctypes.c_int),
        ("egrid_b", ctypes.POINTER(ctypes.c_float)),
        ("bgrid_b", ctypes.POINTER(ctypes.c_float)),
        ("grid_dt", ctypes.c_double),
        ("egrid_d", ctypes.POINTER(ctypes.c_double)),
        ("bgrid_d", ctypes.POINTER(ctypes.c_double)),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.POINTER
        d, i = ctypes.c_double, ctypes.c_int
        L.tpref_set_math.argtypes = [i]
        L.tpref_set_carry_stale.argtypes = [i]
        L.tpref_eval_field.argtypes = [P(TprefField), i, P(d), P(d)]
        L.tpref_wirz_step.argtypes = [P(TprefField), P(d), d, d]
        L.tpref_boris_step.argtypes = [P(TprefField), P(d), d, d]
        L.tpref_diff_coeff.argtypes = [P(TprefField), P(d), P(d), d, P(d)]
        L.tpref_energy_bin.argtypes = [P(d), i, d, d]
        L.tpref_energy_bin.restype = i
        L.tpref_track_fixed.argtypes = [P(TprefField), P(abi.TpTrackParams), i,
                                        P(d), P(d), P(i), P(d), P(d), P(d),
                                        P(i), i, P(d)]
        L.tpref_track_fixed.restype = i
        L.tpref_bs_track.argtypes = [P(TprefField), P(abi.TpBsParams), i, P(d), P(i),
                                     P(ctypes.c_longlong), i, P(d)]
        L.tpref_bs_track.restype = i
        L.tpref_bs_track_diag.argtypes = [P(TprefField), P(abi.TpBsParams),
                                          P(abi.TpTrackParams), i, P(d), P(d), P(i), P(d),
                                          P(d), P(d), P(i), P(ctypes.c_longlong), i, P(d)]
        L.tpref_bs_track_diag.restype = i
        L.tpref_poincare_map.argtypes = [P(TprefField), P(abi.TpPmapParams), i, P(d), P(d),
                                         P(d), P(d), P(i), P(ctypes.c_longlong), i, P(d)]
        L.tpref_poincare_map.restype = i
        L.tpref_sincos.argtypes = [d, P(d), P(d)]
        L.tpref_ellint.argtypes = [d, P(d), P(d)]
        L.tpref_ellint_ref.argtypes = [d, P(d), P(d)]
        L.tpref_set_trajectory.argtypes = [i, P(i), P(d), P(i)]
        L.tpref_mlp_nparams.restype = i
        L.tpref_bf16_round.argtypes = [d]
        L.tpref_bf16_round.restype = d
        L.tpref_gelu.argtypes = [i, P(d), P(d), P(d)]
        L.tpref_gelu_gpu.argtypes = [i, P(d), P(d)]
        L.tpref_mlp_forward.argtypes = [P(d), P(d), i, P(d), P(d), i]
        L.tpref_mlp_loss_grad.argtypes = [P(d), P(d), i, P(d), P(d), P(d), P(d), i]
        _lib = L
    return _lib


class Field:
    """Keeps the numpy buffers a TprefField points into alive."""

    def __init__(self, kind):
        self.s = TprefField()
        self.s.kind = kind
        self._keep = []

    def _hold(self, a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        self._keep.append(a)
        return a

    @classmethod
    def test(cls):
        return cls(FIELD_TEST)

    @classmethod
    def abc(cls, A=1.0, B=np.sqrt(2.0 / 3.0), C=np.sqrt(1.0 / 3.0), lam=1.0,
            B0=1.0):
        f = cls(FIELD_ABC)
        for k, v in enumerate((A, B, C, lam, B0)):
            f.s.abc[k] = v
        return f

    def with_velocity_grid(self, v, lo, hi):
        """v: (nz, ny, nx, 3) float64 velocity nodes."""
        v = self._hold(v, np.float64)
        self.s.vnz, self.s.vny, self.s.vnx = v.shape[:3]
        for k in range(3):
            self.s.vlo[k] = lo[k]
            self.s.vhi[k] = hi[k]
        self.s.vgrid = abi.dptr(v)
        return self

    @classmethod
    def grid(cls, e, b, lo, hi, mode=abi.TP_GATHER_REF_COMPAT):
        """e, b: (nz, ny, nx, 3) float32 nodes (x fastest)."""
        f = cls(FIELD_GRID)
        e = f._hold(e, np.float32)
        b = f._hold(b, np.float32)
        f.s.nz, f.s.ny, f.s.nx = e.shape[:3]
        for k in range(3):
            f.s.lo[k] = lo[k]
            f.s.hi[k] = hi[k]
        f.s.egrid = abi.fptr(e)
        f.s.bgrid = abi.fptr(b)
        f.s.gather_mode = mode
        return f

    @classmethod
    def grid2(cls, e_next, b_next, e_prev, b_prev, lo, hi, grid_dt,
              mode=abi.TP_GATHER_REF_COMPAT):
        """Two time slices: B = dt B_next + (1 - dt) B_prev, dt the fraction
        of t / grid_dt (the reference's multi_tframe = 1)."""
        f = cls.grid(e_next, b_next, lo, hi, mode)
        eb = f._hold(e_prev, np.float32)
        bb = f._hold(b_prev, np.float32)
        f.s.egrid_b = abi.fptr(eb)
        f.s.bgrid_b = abi.fptr(bb)
        f.s.grid_dt = float(grid_dt)
        return f

    @classmethod
    def mhd(cls, v, b, lo, hi):
        """System 3: v, b (nz, ny, nx, 3) float32 velocity and magnetic nodes;
        E = B x v (src/emfields.c:107-120)."""
        f = cls.grid(v, b, lo, hi, abi.TP_GATHER_REF_COMPAT)
        f.s.kind = FIELD_MHD
        return f

    @classmethod
    def grid_double(cls, e, b, lo, hi):
        """e, b: (nz, ny, nx, 3) float64 nodes; the reference's double gathers
        (src/electric_field.c:370-406, src/magnetic_field.c:555-591)."""
        f = cls(FIELD_GRIDD)
        e = f._hold(e, np.float64)
        b = f._hold(b, np.float64)
        f.s.nz, f.s.ny, f.s.nx = e.shape[:3]
        for k in range(3):
            f.s.lo[k] = lo[k]
            f.s.hi[k] = hi[k]
        f.s.egrid_d = abi.dptr(e)
        f.s.bgrid_d = abi.dptr(b)
        f.s.gather_mode = abi.TP_GATHER_REF_COMPAT
        return f

    @classmethod
    def modes(cls, rows, b0z):
        f = cls(FIELD_MODES)
        rows = f._hold(rows, np.float64)
        f.s.nmodes = rows.shape[0]
        f.s.modes = abi.dptr(rows)
        f.s.b0z = b0z
        return f

    @classmethod
    def wlcs(cls, rows, ibessel=0, omega=0.001, norm_i=1.0,
             mode=abi.TP_GATHER_REF_COMPAT):
        f = cls(FIELD_WLCS)
        rows = f._hold(rows, np.float64)
        f.s.nwlcs = rows.shape[0]
        f.s.wlcs_rows = abi.dptr(rows)
        f.s.ibessel = ibessel
        f.s.omega = omega
        f.s.norm_i = norm_i
        f.s.wlcs_mode = mode
        return f


def set_carry_stale(on):
    lib().tpref_set_carry_stale(1 if on else 0)


def set_math(use_libm):
    lib().tpref_set_math(1 if use_libm else 0)


def eval_field(field, pts):
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 4)
    out = np.zeros((pts.shape[0], 6))
    lib().tpref_eval_field(ctypes.byref(field.s), pts.shape[0], abi.dptr(pts),
                           abi.dptr(out))
    return out


def steps(field, ptl, nsteps, dt, qm=1.0, pusher=abi.TP_PUSH_WIRZ,
          params=None):
    """nsteps pushes (+ particle_bc when params is given) per particle."""
    ptl = abi.as_particles(ptl).copy()
    L = lib()
    fn = L.tpref_wirz_step if pusher == abi.TP_PUSH_WIRZ else L.tpref_boris_step
    row = (ctypes.c_double * abi.NFIELD)
    for i in range(ptl.shape[0]):
        buf = row(*ptl[i])
        for _ in range(nsteps):
            fn(ctypes.byref(field.s), buf, dt, qm)
            if params is not None:
                esc = ctypes.c_int(0)
                L.tpref_particle_bc(ctypes.byref(params), buf, ctypes.byref(esc))
        ptl[i] = np.frombuffer(buf, dtype=np.float64)
    return ptl


def diff_coeff(field, p0, p, dt):
    p0 = abi.as_particles(p0)
    p = abi.as_particles(p)
    out = np.zeros((p.shape[0], 8))
    L = lib()
    for i in range(p.shape[0]):
        L.tpref_diff_coeff(ctypes.byref(field.s), abi.dptr(p0[i:i + 1]),
                           abi.dptr(p[i:i + 1]), dt, abi.dptr(out[i:i + 1]))
    return out


def track_fixed(field, params, ptl, ptl_init=None, nthreads=1):
    """Returns dict(ptl, nsteps, espect, espect_escape, dcoeff, remain,
    seconds) -- the raw sums of the reference's fixed-step loop."""
    ptl = abi.as_particles(ptl).copy()
    ptl_init = ptl.copy() if ptl_init is None else abi.as_particles(ptl_init)
    n = ptl.shape[0]
    nd = abi.n_dcoeff_rows(params)
    nsteps = np.zeros(n, dtype=np.int32)
    esp = np.zeros((params.nt_out, params.nbins))
    ese = np.zeros((params.nt_out, params.nbins))
    dco = np.zeros((8, nd))
    rem = np.zeros(nd, dtype=np.int32)
    secs = ctypes.c_double(0.0)
    rc = lib().tpref_track_fixed(ctypes.byref(field.s), ctypes.byref(params), n,
                                 abi.dptr(ptl), abi.dptr(ptl_init),
                                 abi.iptr(nsteps), abi.dptr(esp), abi.dptr(ese),
                                 abi.dptr(dco), abi.iptr(rem), nthreads,
                                 ctypes.byref(secs))
    if rc != 0:
        raise RuntimeError(f"tpref_track_fixed failed: {rc}")
    return dict(ptl=ptl, nsteps=nsteps, espect=esp, espect_escape=ese,
                dcoeff=dco, remain=rem, seconds=secs.value)


def track_fixed_traj(field, params, ptl, nsteps_output, ntraj_accum, ptl_time,
                     ptl_init=None, nthreads=1):
    """track_fixed with trajectory diagnostics; ptl_time (float64, shape
    (ntraj_accum[-1], 11)) holds the caller's initial points and receives the
    others; returns the track_fixed result plus 'ptl_time' and 'points'."""
    ntraj_accum = np.ascontiguousarray(ntraj_accum, dtype=np.int32)
    ptl_time = np.ascontiguousarray(ptl_time, dtype=np.float64).copy()
    points = np.zeros(ntraj_accum.size, dtype=np.int32)
    lib().tpref_set_trajectory(int(nsteps_output), abi.iptr(ntraj_accum), abi.dptr(ptl_time),
                               abi.iptr(points))
    try:
        r = track_fixed(field, params, ptl, ptl_init=ptl_init, nthreads=nthreads)
    finally:
        lib().tpref_set_trajectory(0, None, None, None)
    r["ptl_time"] = ptl_time
    r["points"] = points
    return r


def poincare_map(field, params, starts, nthreads=1):
    """tpref_poincare_map: dict(xpos, ypos, zpos, ncrossed, nsteps, seconds)."""
    st = np.ascontiguousarray(starts, dtype=np.float64).reshape(-1, 3)
    n = st.shape[0]
    ns = params.ncross + 1
    xp, yp, zp = (np.zeros((n, ns)) for _ in range(3))
    nc = np.zeros(n, np.int32)
    nst = np.zeros(n, np.int64)
    sec = ctypes.c_double()
    P = ctypes.POINTER
    rc = lib().tpref_poincare_map(ctypes.byref(field.s), ctypes.byref(params), n,
                                  st.ctypes.data_as(P(ctypes.c_double)),
                                  xp.ctypes.data_as(P(ctypes.c_double)),
                                  yp.ctypes.data_as(P(ctypes.c_double)),
                                  zp.ctypes.data_as(P(ctypes.c_double)),
                                  nc.ctypes.data_as(P(ctypes.c_int)),
                                  nst.ctypes.data_as(P(ctypes.c_longlong)), int(nthreads),
                                  ctypes.byref(sec))
    if rc != 0:
        raise ValueError(f"tpref_poincare_map: {rc}")
    return dict(xpos=xp, ypos=yp, zpos=zp, ncrossed=nc, nsteps=nst, seconds=sec.value)


def bs_track_diag(field, bs, params, ptl, ptl_init=None, nthreads=1):
    """Adaptive Bulirsch-Stoer with dense output over estep*dt, with the
    fixed-step loop's diagnostics at its output times; returns what
    track_fixed returns (nsteps = accepted steps) and the step totals."""
    ptl = abi.as_particles(ptl).copy()
    ptl_init = ptl.copy() if ptl_init is None else abi.as_particles(ptl_init)
    n = ptl.shape[0]
    nd = abi.n_dcoeff_rows(params)
    nsteps = np.zeros(n, dtype=np.int32)
    esp = np.zeros((params.nt_out, params.nbins))
    ese = np.zeros((params.nt_out, params.nbins))
    dco = np.zeros((8, nd))
    rem = np.zeros(nd, dtype=np.int32)
    tot = (ctypes.c_longlong * 4)()
    secs = ctypes.c_double()
    rc = lib().tpref_bs_track_diag(ctypes.byref(field.s), ctypes.byref(bs), ctypes.byref(params),
                                   n, abi.dptr(ptl), abi.dptr(ptl_init), abi.iptr(nsteps),
                                   abi.dptr(esp), abi.dptr(ese), abi.dptr(dco), abi.iptr(rem),
                                   tot, nthreads, ctypes.byref(secs))
    if rc != 0:
        raise RuntimeError(f"tpref_bs_track_diag failed: {rc}")
    return dict(ptl=ptl, nsteps=nsteps, espect=esp, espect_escape=ese, dcoeff=dco,
                remain=rem, accepted=tot[0], rejected=tot[1], rhs=tot[2], failed=tot[3],
                seconds=secs.value)


def bs_track_diag_traj(field, bs, params, ptl, nsteps_output, ntraj_accum, ptl_time,
                       ptl_init=None, nthreads=1):
    """bs_track_diag with trajectory diagnostics (the particle at the
    output steps j = k nsteps_output); ptl_time holds the caller's initial
    points and receives the others; returns the bs_track_diag result plus
    'ptl_time' and 'points'."""
    ntraj_accum = np.ascontiguousarray(ntraj_accum, dtype=np.int32)
    ptl_time = np.ascontiguousarray(ptl_time, dtype=np.float64).copy()
    points = np.zeros(ntraj_accum.size, dtype=np.int32)
    lib().tpref_set_trajectory(int(nsteps_output), abi.iptr(ntraj_accum), abi.dptr(ptl_time),
                               abi.iptr(points))
    try:
        r = bs_track_diag(field, bs, params, ptl, ptl_init=ptl_init, nthreads=nthreads)
    finally:
        lib().tpref_set_trajectory(0, None, None, None)
    r["ptl_time"] = ptl_time
    r["points"] = points
    return r


def bs_track(field, params, ptl, nthreads=1):
    """Adaptive Bulirsch-Stoer of every particle to its end time."""
    ptl = abi.as_particles(ptl).copy()
    n = ptl.shape[0]
    nsteps = np.zeros(n, dtype=np.int32)
    tot = (ctypes.c_longlong * 4)()
    secs = ctypes.c_double()
    rc = lib().tpref_bs_track(ctypes.byref(field.s), ctypes.byref(params), n, abi.dptr(ptl),
                              abi.iptr(nsteps), tot, nthreads, ctypes.byref(secs))
    if rc != 0:
        raise RuntimeError(f"tpref_bs_track failed: {rc}")
    return dict(ptl=ptl, nsteps=nsteps, accepted=tot[0], rejected=tot[1], rhs=tot[2],
                failed=tot[3], seconds=secs.value)


def sincos(x):
    L = lib()
    x = np.asarray(x, dtype=np.float64).ravel()
    s = np.empty_like(x)
    c = np.empty_like(x)
    a, b = ctypes.c_double(), ctypes.c_double()
    for i, v in enumerate(x):
        L.tpref_sincos(float(v), ctypes.byref(a), ctypes.byref(b))
        s[i], c[i] = a.value, b.value
    return s, c


def ellint(k, ref_compat=False):
    a, b = ctypes.c_double(), ctypes.c_double()
    fn = lib().tpref_ellint_ref if ref_compat else lib().tpref_ellint
    fn(float(k), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value   # (E, K)


# --- learned field surrogate (oracle/mlpref.c) -----------------------------------

def mlp_nparams():
    return lib().tpref_mlp_nparams()


def bf16_round(v):
    return np.array([lib().tpref_bf16_round(float(x)) for x in np.ravel(v)])


def gelu_gpu(z):
    """The GPU kernels' activation (FP16 polynomial form), emulated."""
    z = np.ascontiguousarray(z, dtype=np.float64).ravel()
    g = np.empty_like(z)
    lib().tpref_gelu_gpu(z.size, abi.dptr(z), abi.dptr(g))
    return g


def gelu(z):
    z = np.ascontiguousarray(z, dtype=np.float64).ravel()
    g = np.empty_like(z)
    gd = np.empty_like(z)
    lib().tpref_gelu(z.size, abi.dptr(z), abi.dptr(g), abi.dptr(gd))
    return g, gd


def mlp_forward(params, q, x, emulate=False):
    """Network outputs (n, 3) before the output scale and the guide field."""
    params = np.ascontiguousarray(params, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    out = np.empty((x.shape[0], 3))
    lib().tpref_mlp_forward(abi.dptr(params), abi.dptr(q), x.shape[0], abi.dptr(x),
                            abi.dptr(out), 1 if emulate else 0)
    return out


def mlp_loss_grad(params, q, x, y, emulate=False):
    """(mean squared error, gradient over the flat parameter vector)."""
    params = np.ascontiguousarray(params, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1, 3)
    g = np.empty(params.size)
    loss = ctypes.c_double()
    lib().tpref_mlp_loss_grad(abi.dptr(params), abi.dptr(q), x.shape[0], abi.dptr(x),
                              abi.dptr(y), abi.dptr(g), ctypes.byref(loss),
                              1 if emulate else 0)
    return loss.value, g
