"""Python mirror of the C ABI (include/tp_gpu.h) over libtpgpu.so.

Every call goes to the HIP library; there is no CPU fallback.  A missing
library raises ImportError-like RuntimeError at first use.
"""
import ctypes
import os

import numpy as np

from . import abi
from .abi import dptr, fptr, iptr

_lib = None

# tp_plane_reader: (user, slice, comp, s0, s1, dst) -> 0 or nonzero to stop
PLANE_READER = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                ctypes.c_longlong, ctypes.c_longlong,
                                ctypes.POINTER(ctypes.c_float))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(abi.LIB_PATH):
            raise RuntimeError(
                f"{abi.LIB_PATH} is not built; run `python -m test_particle_amd.build`")
        L = ctypes.CDLL(abi.LIB_PATH)
        P = ctypes.POINTER
        d, i, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        L.tp_gpu_init.argtypes = [i, P(vp)]
        L.tp_gpu_finalize.argtypes = [vp]
        L.tp_gpu_finalize.restype = None
        L.tp_gpu_strerror.argtypes = [i]
        L.tp_gpu_strerror.restype = ctypes.c_char_p
        L.tp_gpu_device_info.argtypes = [vp, ctypes.c_char_p, i, P(i)]
        L.tp_gpu_set_field_test.argtypes = [vp]
        L.tp_gpu_set_field_abc.argtypes = [vp, d, d, d, d, d]
        L.tp_gpu_set_velocity_grid.argtypes = [vp, i, i, i, P(d), P(d), P(d)]
        L.tp_gpu_set_field_grid.argtypes = [vp, i, i, i, P(d), P(d), P(ctypes.c_float),
                                            P(ctypes.c_float), i]
        L.tp_gpu_set_field_wlcs.argtypes = [vp, i, P(d), i, d, d, i]
        L.tp_gpu_set_field_grid2.argtypes = [vp, i, i, i, P(d), P(d), P(ctypes.c_float),
                                             P(ctypes.c_float), P(ctypes.c_float),
                                             P(ctypes.c_float), d, i]
        L.tp_gpu_set_field_modes.argtypes = [vp, i, P(d), d]
        L.tp_gpu_set_field_mhd.argtypes = [vp, i, i, i, P(d), P(d), P(ctypes.c_float),
                                           P(ctypes.c_float)]
        L.tp_gpu_set_field_grid_double.argtypes = [vp, i, i, i, P(d), P(d), P(d), P(d)]
        L.tp_gpu_load_field_grid.argtypes = [vp, i, i, i, P(d), P(d), i, i, d, d, d, i,
                                             PLANE_READER, vp, ctypes.c_size_t]
        L.tp_gpu_set_field_grid_planes.argtypes = [vp, i, i, i, P(d), P(d), i, d, d,
                                                   P(ctypes.c_float), P(ctypes.c_float), d, i]
        L.tp_gpu_get_field_nodes.argtypes = [vp, i, i, i, i, P(ctypes.c_float)]
        L.tp_gpu_set_grid_cells.argtypes = [vp, i]
        L.tp_gpu_eval_field.argtypes = [vp, i, P(d), P(d), i]
        L.tp_gpu_track_fixed.argtypes = [vp, P(abi.TpTrackParams), i, P(d), P(d), P(i),
                                         P(d), P(d), P(d), P(i)]
        L.tp_gpu_track_fixed_traj.argtypes = [vp, P(abi.TpTrackParams), i, P(d), P(d), P(i),
                                              P(d), P(d), P(d), P(i), i, P(i), P(d), P(i)]
        L.tp_gpu_load_particles.argtypes = [vp, i, P(d), P(d)]
        L.tp_gpu_run_fixed.argtypes = [vp, P(abi.TpTrackParams)]
        L.tp_gpu_store_particles.argtypes = [vp, P(d), P(i)]
        L.tp_gpu_read_diagnostics.argtypes = [vp, P(d), P(d), P(d), P(i)]
        L.tp_gpu_last_kernel_time.argtypes = [vp, P(d), P(i)]
        L.tp_gpu_run_fixed_noalloc.argtypes = [vp, P(abi.TpTrackParams), vp, vp, vp, vp]
        L.tp_gpu_run_adaptive.argtypes = [vp, P(abi.TpBsParams), P(abi.TpBsStats)]
        L.tp_gpu_track_adaptive.argtypes = [vp, P(abi.TpBsParams), i, P(d), P(i),
                                            P(abi.TpBsStats)]
        L.tp_gpu_tune_adaptive.argtypes = [vp, P(abi.TpBsParams), ctypes.c_char_p, P(i), P(i)]
        L.tp_gpu_run_adaptive_diag.argtypes = [vp, P(abi.TpBsParams), P(abi.TpTrackParams),
                                               P(abi.TpBsStats)]
        L.tp_gpu_track_adaptive_diag_traj.argtypes = [vp, P(abi.TpBsParams),
                                                      P(abi.TpTrackParams), i, P(d), P(d),
                                                      P(i), P(d), P(d), P(d), P(i), i, P(i),
                                                      P(d), P(i), P(abi.TpBsStats)]
        f = ctypes.c_float
        L.tp_gpu_mlp_set.argtypes = [vp, P(f), P(f), d]
        L.tp_gpu_mlp_get.argtypes = [vp, P(f)]
        L.tp_gpu_mlp_forward.argtypes = [vp, i, P(d), P(d)]
        L.tp_gpu_mlp_loss_grad.argtypes = [vp, i, P(f), P(f), P(f), P(d)]
        L.tp_gpu_train_surrogate.argtypes = [vp, P(abi.TpMlpTrainParams),
                                             P(abi.TpMlpTrainStats)]
        L.tp_gpu_set_field_mlp.argtypes = [vp, d, i]
        L.tp_gpu_mlp_fallback_count.argtypes = [vp, P(ctypes.c_longlong)]
        L.tp_gpu_mlp_fallback_flags.argtypes = [vp, i, P(i)]
        L.tp_gpu_poincare_map.argtypes = [vp, P(abi.TpPmapParams), i, P(d), P(d), P(d), P(d),
                                          P(i), P(ctypes.c_longlong)]
        cs = ctypes.c_char_p
        L.tp_host_spectrum_bins.argtypes = [i, d, d, P(d), P(d)]
        L.tp_host_spectrum_t0.argtypes = [i, P(d), i, d, d, P(d)]
        L.tp_host_write_espectrum_t0.argtypes = [cs, cs, i, d, d, P(d), i]
        L.tp_host_append_espectrum.argtypes = [cs, i, i, d, d, P(d)]
        L.tp_host_write_diff_coeffs.argtypes = [cs, i, P(d), P(i)]
        L.tp_host_pmap_starts.argtypes = [i, i, P(d), d, d, d, d, i, i, P(d)]
        L.tp_host_read_pic_gda.argtypes = [cs, i, i, i, i, i, i, d, d, P(f), P(f), P(f), P(f)]
        L.tp_host_load_pic_gda.argtypes = [vp, cs, i, i, i, i, i, i, P(d), P(d), d, d, d, i]
        L.tp_peak_measure.argtypes = [i, i, ctypes.c_size_t, i, P(d), P(d)]
        _lib = L
    return _lib


class TPError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: {lib().tp_gpu_strerror(code).decode()} ({code})")
        self.code = code


def _check(rc, what):
    if rc != 0:
        raise TPError(rc, what)


def measure_peak(what, nbytes=0, reps=5, device=0):
    """tp_peak_measure: (rate, ms) of the fastest of `reps` launches of a
    microbenchmark kernel -- GB/s for abi.TP_PEAK_HBM_*, TFLOP/s for the
    arithmetic ones."""
    v, ms = ctypes.c_double(), ctypes.c_double()
    _check(lib().tp_peak_measure(device, what, nbytes, reps, ctypes.byref(v),
                                 ctypes.byref(ms)), "tp_peak_measure")
    return v.value, ms.value


def _vec3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


class Tracker:
    """One tp_ctx: a field source plus resident particles on one GPU."""

    def __init__(self, device=0):
        self._ctx = ctypes.c_void_p()
        _check(lib().tp_gpu_init(device, ctypes.byref(self._ctx)), "tp_gpu_init")
        self._n = 0

    def close(self):
        if self._ctx:
            lib().tp_gpu_finalize(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_info(self):
        buf = ctypes.create_string_buffer(256)
        cu = ctypes.c_int()
        _check(lib().tp_gpu_device_info(self._ctx, buf, 256, ctypes.byref(cu)), "device_info")
        return buf.value.decode(), cu.value

    # --- fields --------------------------------------------------------
    def set_field_test(self):
        _check(lib().tp_gpu_set_field_test(self._ctx), "set_field_test")

    def set_field_abc(self, A=1.0, B=np.sqrt(2.0 / 3.0), C=np.sqrt(1.0 / 3.0), lam=1.0,
                      B0=1.0):
        _check(lib().tp_gpu_set_field_abc(self._ctx, A, B, C, lam, B0), "set_field_abc")

    def set_velocity_grid(self, v, lo, hi):
        if v is None:
            _check(lib().tp_gpu_set_velocity_grid(self._ctx, 0, 0, 0, None, None, None),
                   "set_velocity_grid")
            return
        v = np.ascontiguousarray(v, dtype=np.float64)
        nz, ny, nx = v.shape[:3]
        _check(lib().tp_gpu_set_velocity_grid(self._ctx, nx, ny, nz, _vec3(lo), _vec3(hi),
                                              dptr(v)), "set_velocity_grid")

    def set_field_grid(self, e, b, lo, hi, mode=abi.TP_GATHER_REF_COMPAT):
        e = np.ascontiguousarray(e, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        nz, ny, nx = e.shape[:3]
        _check(lib().tp_gpu_set_field_grid(self._ctx, nx, ny, nz, _vec3(lo), _vec3(hi),
                                           fptr(e), fptr(b), mode), "set_field_grid")

    def set_field_mhd(self, v, b, lo, hi):
        """System 3: float velocity and magnetic nodes (nz, ny, nx, 3); E = B x v
        (tp_gpu_set_field_mhd)."""
        v = np.ascontiguousarray(v, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        if v.shape != b.shape:
            raise ValueError("v and b must have the same shape")
        nz, ny, nx = v.shape[:3]
        _check(lib().tp_gpu_set_field_mhd(self._ctx, nx, ny, nz, _vec3(lo), _vec3(hi),
                                          fptr(v), fptr(b)), "set_field_mhd")

    def set_field_grid_double(self, e, b, lo, hi):
        """PIC grid of doubles (nz, ny, nx, 3) (tp_gpu_set_field_grid_double)."""
        e = np.ascontiguousarray(e, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        if e.shape != b.shape:
            raise ValueError("e and b must have the same shape")
        nz, ny, nx = e.shape[:3]
        _check(lib().tp_gpu_set_field_grid_double(self._ctx, nx, ny, nz, _vec3(lo),
                                                  _vec3(hi), dptr(e), dptr(b)),
               "set_field_grid_double")

    def set_field_grid2(self, e_next, b_next, e_prev, b_prev, lo, hi, grid_dt,
                        mode=abi.TP_GATHER_REF_COMPAT):
        """Two time slices, blended dt * next + (1 - dt) * previous."""
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (e_next, b_next, e_prev,
                                                                     b_prev)]
        if any(a.shape != arrs[0].shape for a in arrs):
            raise ValueError("the four slices must have the same shape")
        nz, ny, nx = arrs[0].shape[:3]
        _check(lib().tp_gpu_set_field_grid2(self._ctx, nx, ny, nz, _vec3(lo), _vec3(hi),
                                            *(fptr(a) for a in arrs), float(grid_dt), mode),
               "set_field_grid2")

    def set_grid_cells(self, mode):
        """Cell-major copy of the PIC grids set after this call
        (tp_gpu_set_grid_cells): abi.TP_CELLS_AUTO, _OFF or _ON."""
        _check(lib().tp_gpu_set_grid_cells(self._ctx, int(mode)), "set_grid_cells")

    def set_field_grid_planes(self, planes_next, lo, hi, order=abi.TP_PLANES_X_FASTEST,
                              e0=1.0, b0=1.0, planes_prev=None, grid_dt=0.0,
                              mode=abi.TP_GATHER_REF_COMPAT):
        """PIC field from raw component planes (tp_gpu_set_field_grid_planes):
        planes_* of shape (6, nz, ny, nx) for TP_PLANES_X_FASTEST or
        (6, nx, ny, nz) for TP_PLANES_X_SLOWEST, components ex ey ez bx by bz,
        unscaled; values become (float)(v * e0) and (float)(v * b0)."""
        pn = np.ascontiguousarray(planes_next, dtype=np.float32)
        if pn.ndim != 4 or pn.shape[0] != 6:
            raise ValueError("planes must have shape (6, n0, n1, n2)")
        pp = None if planes_prev is None else np.ascontiguousarray(planes_prev,
                                                                   dtype=np.float32)
        if pp is not None and pp.shape != pn.shape:
            raise ValueError("the two slices must have the same shape")
        a, bb, c = pn.shape[1:]
        nz, ny, nx = (a, bb, c) if order == abi.TP_PLANES_X_FASTEST else (c, bb, a)
        _check(lib().tp_gpu_set_field_grid_planes(self._ctx, nx, ny, nz, _vec3(lo), _vec3(hi),
                                                  int(order), float(e0), float(b0), fptr(pn),
                                                  None if pp is None else fptr(pp),
                                                  float(grid_dt), mode), "set_field_grid_planes")

    def load_field_grid(self, reader, shape, lo, hi, order=abi.TP_PLANES_X_FASTEST,
                        two_slices=False, e0=1.0, b0=1.0, grid_dt=0.0,
                        mode=abi.TP_GATHER_REF_COMPAT, stage_bytes=0):
        """PIC field through a plane reader (tp_gpu_load_field_grid): reader
        (slice, comp, s0, s1) returns the float32 values of that slab in the
        plane's order, or None to stop the load; shape = (nx, ny, nz)."""
        nx, ny, nz = (int(v) for v in shape)

        want = (ny * nx if order == abi.TP_PLANES_X_FASTEST else ny * nz)

        def cb(user, slice_, comp, s0, s1, dst):
            try:
                v = reader(slice_, comp, s0, s1)
                if v is None:
                    return 1
                v = np.ascontiguousarray(v, dtype=np.float32).ravel()
                if v.size != (s1 - s0) * want:
                    return 1
                ctypes.memmove(dst, v.ctypes.data, v.nbytes)
                return 0
            except Exception:   # an exception must not cross the C frames
                return 1

        fn = PLANE_READER(cb)
        _check(lib().tp_gpu_load_field_grid(self._ctx, nx, ny, nz, _vec3(lo), _vec3(hi),
                                            int(order), 1 if two_slices else 0, float(e0),
                                            float(b0), float(grid_dt), mode, fn, None,
                                            int(stage_bytes)), "load_field_grid")

    def load_pic_gda(self, path, ct, shape, lo, hi, e0, b0, nt=0, two_slices=False,
                     grid_dt=0.0, mode=abi.TP_GATHER_REF_COMPAT):
        """The reference's .gda field files straight onto the GPU
        (tp_host_load_pic_gda); shape = (nx, ny, nz)."""
        nx, ny, nz = (int(v) for v in shape)
        p = path if path.endswith(os.sep) else path + os.sep
        _check(lib().tp_host_load_pic_gda(self._ctx, p.encode(), int(ct), int(nt),
                                          1 if two_slices else 0, nx, ny, nz, _vec3(lo),
                                          _vec3(hi), float(e0), float(b0), float(grid_dt),
                                          mode), "load_pic_gda")

    def field_nodes(self, shape, slice_=0):
        """The grid field's nodes as the gathers read them
        (tp_gpu_get_field_nodes); shape = (nx, ny, nz) of the grid.  Returns
        float32 (nz, ny, nx, 8): ex ey ez bx by bz 0 0."""
        nx, ny, nz = (int(v) for v in shape)
        out = np.zeros((nz, ny, nx, 8), np.float32)
        _check(lib().tp_gpu_get_field_nodes(self._ctx, int(slice_), nx, ny, nz, fptr(out)),
               "field_nodes")
        return out

    def set_field_wlcs(self, rows, ibessel=0, omega=0.001, norm_i=1.0,
                       mode=abi.TP_GATHER_REF_COMPAT):
        rows = np.ascontiguousarray(rows, dtype=np.float64).reshape(-1, 14)
        _check(lib().tp_gpu_set_field_wlcs(self._ctx, rows.shape[0], dptr(rows), ibessel,
                                           omega, norm_i, mode), "set_field_wlcs")

    def set_field_modes(self, rows, b0z):
        rows = np.ascontiguousarray(rows, dtype=np.float64).reshape(-1, 7)
        _check(lib().tp_gpu_set_field_modes(self._ctx, rows.shape[0], dptr(rows), b0z),
               "set_field_modes")

    def eval_field(self, pts, precision=abi.TP_FP64):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 4)
        out = np.zeros((pts.shape[0], 6))
        _check(lib().tp_gpu_eval_field(self._ctx, pts.shape[0], dptr(pts), dptr(out),
                                       precision), "eval_field")
        return out

    # --- tracking --------------------------------------------------------
    def track_fixed(self, params, ptl, ptl_init=None):
        """The reference's fixed-step loop (src/tracking.c:90-190) on the
        GPU; returns the same raw arrays as the oracle."""
        ptl = abi.as_particles(ptl).copy()
        ptl_init = ptl.copy() if ptl_init is None else abi.as_particles(ptl_init)
        n = ptl.shape[0]
        nd = abi.n_dcoeff_rows(params)
        nsteps = np.zeros(n, dtype=np.int32)
        esp = np.zeros((params.nt_out, params.nbins))
        ese = np.zeros((params.nt_out, params.nbins))
        dco = np.zeros((8, nd))
        rem = np.zeros(nd, dtype=np.int32)
        _check(lib().tp_gpu_track_fixed(self._ctx, ctypes.byref(params), n, dptr(ptl),
                                        dptr(ptl_init), iptr(nsteps), dptr(esp), dptr(ese),
                                        dptr(dco), iptr(rem)), "track_fixed")
        self._n = n
        return dict(ptl=ptl, nsteps=nsteps, espect=esp, espect_escape=ese, dcoeff=dco,
                    remain=rem)

    def track_fixed_traj(self, params, ptl, nsteps_output, ntraj_accum, ptl_time,
                         ptl_init=None):
        """track_fixed with the reference's trajectory diagnostics
        (traj_diagnose = 1); ptl_time holds the initial points on entry."""
        ptl = abi.as_particles(ptl).copy()
        ptl_init = ptl.copy() if ptl_init is None else abi.as_particles(ptl_init)
        n = ptl.shape[0]
        nd = abi.n_dcoeff_rows(params)
        acc = np.ascontiguousarray(ntraj_accum, dtype=np.int32)
        pt = abi.as_particles(ptl_time).copy()
        if acc.size != n or (n and pt.shape[0] < acc[-1]):
            raise ValueError("ntraj_accum must have n entries and ptl_time ntraj_accum[-1] rows")
        nsteps = np.zeros(n, dtype=np.int32)
        points = np.zeros(n, dtype=np.int32)
        esp = np.zeros((params.nt_out, params.nbins))
        ese = np.zeros((params.nt_out, params.nbins))
        dco = np.zeros((8, nd))
        rem = np.zeros(nd, dtype=np.int32)
        _check(lib().tp_gpu_track_fixed_traj(self._ctx, ctypes.byref(params), n, dptr(ptl),
                                             dptr(ptl_init), iptr(nsteps), dptr(esp), dptr(ese),
                                             dptr(dco), iptr(rem), int(nsteps_output),
                                             iptr(acc), dptr(pt), iptr(points)),
               "track_fixed_traj")
        return dict(ptl=ptl, nsteps=nsteps, espect=esp, espect_escape=ese, dcoeff=dco,
                    remain=rem, ptl_time=pt, points=points)

    def load(self, ptl, ptl_init=None):
        ptl = abi.as_particles(ptl)
        ptl_init = ptl if ptl_init is None else abi.as_particles(ptl_init)
        _check(lib().tp_gpu_load_particles(self._ctx, ptl.shape[0], dptr(ptl),
                                           dptr(ptl_init)), "load_particles")
        self._n = ptl.shape[0]

    def run(self, params):
        _check(lib().tp_gpu_run_fixed(self._ctx, ctypes.byref(params)), "run_fixed")

    def run_noalloc(self, params, espect, espect_escape, dcoeff, remain):
        """tp_gpu_run_fixed_noalloc into caller-owned DEVICE buffers given as
        integer addresses: counts [nt_out][nbins] uint32 (twice), D sums [8][nd]
        float64, remaining counts [nd] int32."""
        rc = lib().tp_gpu_run_fixed_noalloc(self._ctx, ctypes.byref(params),
                                            ctypes.c_void_p(espect),
                                            ctypes.c_void_p(espect_escape),
                                            ctypes.c_void_p(dcoeff), ctypes.c_void_p(remain))
        _check(rc, "run_fixed_noalloc")

    def store(self):
        ptl = np.zeros((self._n, abi.NFIELD))
        nsteps = np.zeros(self._n, dtype=np.int32)
        _check(lib().tp_gpu_store_particles(self._ctx, dptr(ptl), iptr(nsteps)),
               "store_particles")
        return ptl, nsteps

    def diagnostics(self, params):
        nd = abi.n_dcoeff_rows(params)
        esp = np.zeros((params.nt_out, params.nbins))
        ese = np.zeros((params.nt_out, params.nbins))
        dco = np.zeros((8, nd))
        rem = np.zeros(nd, dtype=np.int32)
        _check(lib().tp_gpu_read_diagnostics(self._ctx, dptr(esp), dptr(ese), dptr(dco),
                                             iptr(rem)), "read_diagnostics")
        return dict(espect=esp, espect_escape=ese, dcoeff=dco, remain=rem)

    # --- adaptive Bulirsch-Stoer ------------------------------------------
    def run_adaptive(self, params):
        st = abi.TpBsStats()
        _check(lib().tp_gpu_run_adaptive(self._ctx, ctypes.byref(params), ctypes.byref(st)),
               "run_adaptive")
        return dict(accepted=st.accepted, rejected=st.rejected, rhs=st.rhs_evals,
                    failed=st.failed, kernel_ms=st.kernel_ms)

    def run_adaptive_diag(self, bs, params):
        """tp_gpu_run_adaptive_diag on the resident particles: the adaptive
        stepper with dense output over estep*dt and the fixed-step loop's
        diagnostics (read with diagnostics(params))."""
        st = abi.TpBsStats()
        _check(lib().tp_gpu_run_adaptive_diag(self._ctx, ctypes.byref(bs), ctypes.byref(params),
                                              ctypes.byref(st)), "run_adaptive_diag")
        return dict(accepted=st.accepted, rejected=st.rejected, rhs=st.rhs_evals,
                    failed=st.failed, kernel_ms=st.kernel_ms)

    def track_adaptive_diag(self, bs, params, ptl, ptl_init=None):
        """load + run_adaptive_diag + store + diagnostics; returns what
        track_fixed returns (nsteps = accepted steps) and the step totals."""
        self.load(ptl, ptl_init)
        r = self.run_adaptive_diag(bs, params)
        p, ns = self.store()
        r.update(ptl=p, nsteps=ns, **self.diagnostics(params))
        return r

    def track_adaptive_diag_traj(self, bs, params, ptl, nsteps_output, ntraj_accum, ptl_time,
                                 ptl_init=None):
        """track_adaptive_diag with the reference's trajectory diagnostics
        (tracking_method = 1, traj_diagnose = 1); ptl_time holds the initial
        points on entry."""
        ptl = abi.as_particles(ptl).copy()
        ptl_init = ptl.copy() if ptl_init is None else abi.as_particles(ptl_init)
        n = ptl.shape[0]
        nd = abi.n_dcoeff_rows(params)
        acc = np.ascontiguousarray(ntraj_accum, dtype=np.int32)
        pt = abi.as_particles(ptl_time).copy()
        if acc.size != n or (n and pt.shape[0] < acc[-1]):
            raise ValueError("ntraj_accum must have n entries and ptl_time ntraj_accum[-1] rows")
        nsteps = np.zeros(n, dtype=np.int32)
        points = np.zeros(n, dtype=np.int32)
        esp = np.zeros((params.nt_out, params.nbins))
        ese = np.zeros((params.nt_out, params.nbins))
        dco = np.zeros((8, nd))
        rem = np.zeros(nd, dtype=np.int32)
        st = abi.TpBsStats()
        _check(lib().tp_gpu_track_adaptive_diag_traj(
            self._ctx, ctypes.byref(bs), ctypes.byref(params), n, dptr(ptl), dptr(ptl_init),
            iptr(nsteps), dptr(esp), dptr(ese), dptr(dco), iptr(rem), int(nsteps_output),
            iptr(acc), dptr(pt), iptr(points), ctypes.byref(st)), "track_adaptive_diag_traj")
        return dict(ptl=ptl, nsteps=nsteps, espect=esp, espect_escape=ese, dcoeff=dco,
                    remain=rem, ptl_time=pt, points=points, accepted=st.accepted,
                    rejected=st.rejected, rhs=st.rhs_evals, failed=st.failed,
                    kernel_ms=st.kernel_ms)

    def track_adaptive(self, params, ptl):
        ptl = abi.as_particles(ptl).copy()
        nsteps = np.zeros(ptl.shape[0], dtype=np.int32)
        st = abi.TpBsStats()
        _check(lib().tp_gpu_track_adaptive(self._ctx, ctypes.byref(params), ptl.shape[0],
                                           dptr(ptl), iptr(nsteps), ctypes.byref(st)),
               "track_adaptive")
        return dict(ptl=ptl, nsteps=nsteps, accepted=st.accepted, rejected=st.rejected,
                    rhs=st.rhs_evals, failed=st.failed, kernel_ms=st.kernel_ms)

    def tune_adaptive(self, params, cache_path=None):
        b = ctypes.c_int()
        w = ctypes.c_int()
        cp = cache_path.encode() if cache_path else None
        _check(lib().tp_gpu_tune_adaptive(self._ctx, ctypes.byref(params), cp, ctypes.byref(b),
                                          ctypes.byref(w)), "tune_adaptive")
        return b.value, w.value

    # --- learned field surrogate ---------------------------------------------
    def mlp_set(self, q, params, scale):
        q = np.ascontiguousarray(q, dtype=np.float32).reshape(abi.TP_MLP_NFREQ, 3)
        params = np.ascontiguousarray(params, dtype=np.float32).ravel()
        if params.size != abi.TP_MLP_NPARAMS:
            raise ValueError(f"expected {abi.TP_MLP_NPARAMS} parameters")
        _check(lib().tp_gpu_mlp_set(self._ctx, fptr(q), fptr(params), float(scale)), "mlp_set")

    def mlp_get(self):
        out = np.zeros(abi.TP_MLP_NPARAMS, dtype=np.float32)
        _check(lib().tp_gpu_mlp_get(self._ctx, fptr(out)), "mlp_get")
        return out

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        out = np.zeros_like(x)
        _check(lib().tp_gpu_mlp_forward(self._ctx, x.shape[0], dptr(x), dptr(out)),
               "mlp_forward")
        return out

    def mlp_loss_grad(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
        y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1, 3)
        g = np.zeros(abi.TP_MLP_NPARAMS, dtype=np.float32)
        loss = ctypes.c_double()
        _check(lib().tp_gpu_mlp_loss_grad(self._ctx, x.shape[0], fptr(x), fptr(y), fptr(g),
                                          ctypes.byref(loss)), "mlp_loss_grad")
        return loss.value, g

    def train_surrogate(self, params):
        st = abi.TpMlpTrainStats()
        _check(lib().tp_gpu_train_surrogate(self._ctx, ctypes.byref(params), ctypes.byref(st)),
               "train_surrogate")
        return dict(first_loss=st.first_loss, last_loss=st.last_loss, steps=st.steps,
                    sample_ms=st.sample_ms, train_ms=st.train_ms)

    def set_field_mlp(self, tol=1e-3, check_interval=100):
        _check(lib().tp_gpu_set_field_mlp(self._ctx, float(tol), int(check_interval)),
               "set_field_mlp")

    def mlp_fallback_count(self):
        v = ctypes.c_longlong()
        _check(lib().tp_gpu_mlp_fallback_count(self._ctx, ctypes.byref(v)),
               "mlp_fallback_count")
        return v.value

    def mlp_fallback_flags(self):
        """Per resident particle: 1 if it used the exact field at the end of
        the last surrogate run."""
        f = np.zeros(self._n, dtype=np.int32)
        _check(lib().tp_gpu_mlp_fallback_flags(self._ctx, self._n, iptr(f)),
               "mlp_fallback_flags")
        return f

    def poincare_map(self, params, starts):
        """tp_gpu_poincare_map: returns dict(xpos, ypos, zpos) of shape
        (npoints, ncross + 1), ncrossed (npoints,) and nsteps (npoints,)."""
        st = np.ascontiguousarray(starts, dtype=np.float64).reshape(-1, 3)
        n = st.shape[0]
        ns = params.ncross + 1
        xp, yp, zp = (np.zeros((n, ns)) for _ in range(3))
        nc = np.zeros(n, np.int32)
        nst = np.zeros(n, np.int64)
        _check(lib().tp_gpu_poincare_map(
            self._ctx, ctypes.byref(params), n, dptr(st), dptr(xp), dptr(yp), dptr(zp),
            iptr(nc), nst.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))), "poincare_map")
        return dict(xpos=xp, ypos=yp, zpos=zp, ncrossed=nc, nsteps=nst)

    def last_kernel_time(self):
        ms = ctypes.c_double()
        k = ctypes.c_int()
        _check(lib().tp_gpu_last_kernel_time(self._ctx, ctypes.byref(ms), ctypes.byref(k)),
               "last_kernel_time")
        return ms.value, k.value


# --- host-side writers (include/tp_host.h) -----------------------------------
def spectrum_bins(nbins, emin, emax):
    eb = np.zeros(nbins)
    ei = np.zeros(nbins)
    _check(lib().tp_host_spectrum_bins(nbins, emin, emax, dptr(eb), dptr(ei)), "spectrum_bins")
    return eb, ei


def spectrum_t0(ptl, nbins, emin, emax):
    ptl = abi.as_particles(ptl)
    c = np.zeros(nbins)
    _check(lib().tp_host_spectrum_t0(ptl.shape[0], dptr(ptl), nbins, emin, emax, dptr(c)),
           "spectrum_t0")
    return c


def write_outputs(dirname, ptl_init, result, nbins, emin, emax, bc_flag=0):
    """The files the reference writes for one run: data/espectrum.dat
    (t = 0 spectrum then one row per frame), data/espect_escape.dat (open
    boundaries) and data/diff_coeffs.dat."""
    import os
    os.makedirs(dirname, exist_ok=True)
    L = lib()
    f_sp = os.path.join(dirname, "espectrum.dat").encode()
    f_es = os.path.join(dirname, "espect_escape.dat").encode()
    t0 = spectrum_t0(ptl_init, nbins, emin, emax)
    _check(L.tp_host_write_espectrum_t0(f_sp, f_es, nbins, emin, emax, dptr(t0), bc_flag),
           "write_espectrum_t0")
    esp = np.ascontiguousarray(result["espect"], dtype=np.float64)
    _check(L.tp_host_append_espectrum(f_sp, nbins, esp.shape[0], emin, emax, dptr(esp)),
           "append_espectrum")
    if bc_flag == 1:
        ese = np.ascontiguousarray(result["espect_escape"], dtype=np.float64)
        _check(L.tp_host_append_espectrum(f_es, nbins, ese.shape[0], emin, emax, dptr(ese)),
               "append_espectrum")
    d = np.ascontiguousarray(result["dcoeff"], dtype=np.float64)
    r = np.ascontiguousarray(result["remain"], dtype=np.int32)
    _check(L.tp_host_write_diff_coeffs(os.path.join(dirname, "diff_coeffs.dat").encode(),
                                       d.shape[1], dptr(d), iptr(r)), "write_diff_coeffs")


def read_pic_gda(path, ct, nx, ny, nz, e0, b0, nt=None, two_slices=False):
    """The reference's PIC field files (include/tp_host.h,
    tp_host_read_pic_gda): returns (e, b) of slice ct, or with two_slices
    (e_next, b_next, e_prev, b_prev), each float32 (nz, ny, nx, 3)."""
    shape = (nz, ny, nx, 3)
    en, bn = np.zeros(shape, np.float32), np.zeros(shape, np.float32)
    ep = np.zeros(shape, np.float32) if two_slices else None
    bp = np.zeros(shape, np.float32) if two_slices else None
    p = path if path.endswith(os.sep) else path + os.sep
    _check(lib().tp_host_read_pic_gda(p.encode(), int(ct), int(nt or 0), 1 if two_slices else 0,
                                      nx, ny, nz, float(e0), float(b0), fptr(en), fptr(bn),
                                      fptr(ep), fptr(bp)), "read_pic_gda")
    return (en, bn, ep, bp) if two_slices else (en, bn)


def pmap_starts(face, p0, single=False, min_h=0.0, max_h=0.0, min_v=0.0, max_v=0.0, nh=1,
                nv=1):
    """Start points of a Poincare map (include/tp_host.h, tp_host_pmap_starts):
    (npoints, 3) float64, point iv*nh + ih at (min_h + ih dh, min_v + iv dv)
    on the two axes other than face."""
    n = 1 if single else int(nh) * int(nv)
    out = np.zeros((n, 3))
    p = np.ascontiguousarray(p0, dtype=np.float64)
    rc = lib().tp_host_pmap_starts(int(face), 1 if single else 0, dptr(p), float(min_h),
                                   float(max_h), float(min_v), float(max_v), int(nh), int(nv),
                                   dptr(out))
    if rc < 0:
        _check(rc, "pmap_starts")
    return out
