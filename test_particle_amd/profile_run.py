"""Fixed workloads for profiling runs (rocprofv3) of the tracking kernels.

    python -m test_particle_amd.profile_run grid  [--steps 1000]   # PIC-grid tracking
    python -m test_particle_amd.profile_run abc   [--steps 1000]   # ABC field tracking
    python -m test_particle_amd.profile_run copy  [--mib 1024]     # plain copy kernel
    python -m test_particle_amd.profile_run pmap  [--side 512]     # Poincare map, PIC grid
    python -m test_particle_amd.profile_run fieldload [--side 384] [--h5 DIR]  # PIC field input
    python -m test_particle_amd.profile_run sursweep [--epochs 4]  # surrogate residual vs
                                                                # its Fourier frequencies

``copy`` launches one device-to-device copy kernel (torch's) of a known size;
its FETCH_SIZE / WRITE_SIZE readings calibrate those counters against bytes.
"""
import argparse
import sys
import time

import numpy as np


def field_load(tr, n, h5dir):
    """A side^3 PIC grid from raw planes: the GPU pipeline (pinned slabs,
    transpose and interleave on the GPU) against the host path of the
    reference's loaders (scale and transpose on the host, then
    tp_gpu_set_field_grid)."""
    from . import abi
    rng = np.random.default_rng(1)
    lo, hi = (0.0, 0.0, 0.0), (float(n), float(n), float(n))
    planes = rng.standard_normal((6, n, n, n), dtype=np.float32)   # x slowest
    gb = planes.nbytes / 1e9
    print(f"fieldload: {n}^3 nodes, {gb:.2f} GB of raw planes", flush=True)
    for order, name in ((abi.TP_PLANES_X_SLOWEST, "x-slowest (HDF5)"),
                        (abi.TP_PLANES_X_FASTEST, "x-fastest (.gda)")):
        for rep in range(2):
            t0 = time.perf_counter()
            tr.set_field_grid_planes(planes, lo, hi, order=order, e0=0.3, b0=1.7)
            w = time.perf_counter() - t0
            print(f"  planes {name} #{rep}: {w * 1e3:.1f} ms, {gb / w:.2f} GB/s", flush=True)
    t0 = time.perf_counter()
    p = planes.astype(np.float64)
    e = np.ascontiguousarray((p[:3] * 0.3).astype(np.float32).transpose(3, 2, 1, 0))
    b = np.ascontiguousarray((p[3:] * 1.7).astype(np.float32).transpose(3, 2, 1, 0))
    del p
    t1 = time.perf_counter()
    tr.set_field_grid(e, b, lo, hi)
    t2 = time.perf_counter()
    print(f"  host path: scale+transpose (numpy) {(t1 - t0) * 1e3:.1f} ms, "
          f"tp_gpu_set_field_grid {(t2 - t1) * 1e3:.1f} ms, {gb / (t2 - t0):.2f} GB/s",
          flush=True)
    if h5dir:
        from . import h5
        fname = f"{h5dir}/fields_0.h5"
        h5.write_pic_planes(fname, "Timestep_0", planes)
        for rep in range(2):
            t0 = time.perf_counter()
            h5.load_pic(tr, fname, "Timestep_0", (n, n, n), lo, hi, 0.3, 1.7)
            w = time.perf_counter() - t0
            print(f"  h5 load_pic #{rep}: {w * 1e3:.1f} ms, {gb / w:.2f} GB/s", flush=True)
        t0 = time.perf_counter()
        e2, b2 = h5.read_pic(fname, "Timestep_0", (n, n, n), 0.3, 1.7)
        t1 = time.perf_counter()
        tr.set_field_grid(e2, b2, lo, hi)
        t2 = time.perf_counter()
        print(f"  h5 host path: tp_h5_read_pic {(t1 - t0) * 1e3:.1f} ms, "
              f"tp_gpu_set_field_grid {(t2 - t1) * 1e3:.1f} ms", flush=True)
        assert np.array_equal(e2, e) and np.array_equal(b2, b)


def sur_sweep(tr, epochs, seed=20261202):
    """The surrogate of the 1000-mode turbulence (the bench's training: 2^24
    samples, batch 16384) trained from the same initial weights with
    Fourier frequencies of several spreads -- Gaussian lattice vectors of
    standard deviation s, and radii log-uniform up to rmax -- and its
    residual |B_net - B| / |B| at 65,536 uniform points."""
    from . import abi, surrogate
    rows, period = surrogate.surrogate_field(1000, seed=seed)
    scale = surrogate.output_scale(rows)
    for cut in (8, 12, 16, 32, 64):
        print(f"field beyond |n| = {cut}: rms {surrogate.spectral_tail(rows, period, cut):.2e} "
              f"of |B|", flush=True)
    xs = np.random.default_rng(seed + 1).uniform(0.0, period, (1 << 16, 3))
    pts = np.zeros((xs.shape[0], 4))
    pts[:, :3] = xs
    tr.set_field_modes(rows, 1.0)
    exact = tr.eval_field(pts)[:, :3]
    cases = [(f"gauss s={s:g}", surrogate.fourier_frequencies(period, sigma=s, seed=seed))
             for s in (2.0, 3.0, 4.0, 6.0, 8.0, 12.0)]
    cases += [(f"logr rmax={r:g}",
               surrogate.fourier_frequencies_logradial(period, r, seed=seed))
              for r in (8.0, 16.0, 32.0, 64.0)]
    for name, q in cases:
        reach = np.max(np.linalg.norm(q.astype(np.float64) * period, axis=1))
        tr.mlp_set(q, surrogate.init_params(seed), scale)
        st = tr.train_surrogate(abi.make_mlp_train_params(
            nsamples=1 << 24, period=period, batch=16384, epochs=epochs, lr=2e-3,
            lr_min=2e-5, seed=seed))
        bnet = scale * tr.mlp_forward(xs)
        bnet[:, 2] += 1.0
        res = np.linalg.norm(bnet - exact, axis=1) / np.linalg.norm(exact, axis=1)
        print(f"epochs {epochs:2d} {name:15s} max |n| {reach:5.1f} loss {st['last_loss']:.3e} "
              f"median {np.median(res):.2e} p90 {np.quantile(res, .9):.2e} "
              f"p99 {np.quantile(res, .99):.2e} below 1e-3 {np.mean(res < 1e-3):.3f}",
              flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["grid", "abc", "copy", "modes", "surrogate", "pmap",
                                       "fieldload", "sursweep"])
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--particles", type=int, default=1 << 20)
    ap.add_argument("--pusher", default="boris", choices=["boris", "wirz"])
    ap.add_argument("--fp32", action="store_true")
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--nmodes", type=int, default=1000)
    ap.add_argument("--tgyro", type=float, default=0.05, help="adaptive horizon (gyroperiods)")
    ap.add_argument("--tune", action="store_true")
    ap.add_argument("--segments", type=int, default=1,
                    help="modes: runs of --tgyro gyroperiods each, one after the other")
    ap.add_argument("--dense", action="store_true",
                    help="modes: dense output with the fixed-step diagnostics (--steps outputs)")
    ap.add_argument("--traj", type=int, default=0,
                    help="modes --dense: also record every particle every TRAJ output steps "
                         "(tp_gpu_track_adaptive_diag_traj)")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--wpc", type=int, default=0)
    ap.add_argument("--tol", type=float, default=1e-3, help="surrogate residual tolerance")
    ap.add_argument("--epochs", type=int, default=0, help="surrogate training epochs")
    ap.add_argument("--exact", action="store_true", help="surrogate workload, exact field")
    ap.add_argument("--side", type=int, default=512, help="Poincare start points per side")
    ap.add_argument("--ncross", type=int, default=2, help="Poincare crossings per line")
    ap.add_argument("--chunk", type=int, default=0, help="Poincare RK4 steps per launch")
    ap.add_argument("--lanes", type=int, default=0, choices=[0, 1, 4],
                    help="Poincare lanes per field line (0 automatic)")
    ap.add_argument("--h5", default="", help="fieldload: also through an HDF5 file in DIR")
    ap.add_argument("--two", action="store_true", help="grid: two time slices (dt 20)")
    ap.add_argument("--cells", default="auto", choices=["auto", "off", "on"],
                    help="grid: cell-major copy of the grid")
    ap.add_argument("--dinterval", type=int, default=100,
                    help="grid/abc: steps between diffusion rows (a launch ends at each)")
    ap.add_argument("--sort", action="store_true",
                    help="grid: order the particles by the cell of their start point")
    ap.add_argument("--reps", type=int, default=1,
                    help="grid/abc: runs of --steps steps, each from the initial particles")
    a = ap.parse_args()
    if a.what == "copy":
        import torch
        n = a.mib * (1 << 18)
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copy: {a.mib} MiB", flush=True)
        return
    from . import abi
    from .fields import abc_b_ref, gyroperiod, pic_turbulence_grid
    from .gpu import Tracker
    from .particles import init_particles
    tr = Tracker(0)
    if a.what == "modes":
        from .fields import turbulence_modes
        rows = turbulence_modes(a.nmodes)
        tr.set_field_modes(rows, 1.0)
        ptl = init_particles(a.particles, (0, 0, 0), (100, 100, 100), 0.1, kind="loguniform",
                             vmax=0.9)
        tr.load(ptl)
        prm = abi.make_bs_params(rtol=1e-8, atol=1e-10, t_end_gyro=a.tgyro, b_ref=1.0,
                                 block=a.block, waves_per_cu=a.wpc)
        if a.tune:
            t0 = time.perf_counter()
            b, w = tr.tune_adaptive(prm)
            print(f"tuned: block {b}, waves/CU {w} ({time.perf_counter() - t0:.1f}s)", flush=True)
            tr.load(ptl)
        if a.dense:
            # dense output with the fixed-step loop's diagnostics: output
            # step 0.1 over 200 steps, D rows every 20, no boundary
            dp = abi.make_params(dt=0.1, estep=a.steps, lo=(0, 0, 0), hi=(100, 100, 100),
                                 pusher=abi.TP_PUSH_BORIS, bc_flag=abi.TP_BC_NONE, nbins=102,
                                 nt_out=10, emin=1e-4, emax=1.0, dcoeff_interval=20)
            if a.traj > 0:
                # every particle's records, slot 0 the initial point (the
                # load, store and copies are inside the wall time here)
                import numpy as np
                slots = a.steps // a.traj + 1
                accum = (np.arange(1, a.particles + 1, dtype=np.int64) * slots).astype(np.int32)
                pt = np.zeros((a.particles * slots, abi.NFIELD))
                pt[::slots] = ptl
                t0 = time.perf_counter()
                st = tr.track_adaptive_diag_traj(prm, dp, ptl, a.traj, accum, pt)
                wall = time.perf_counter() - t0
                print(f"modes({a.nmodes}) dense + records every {a.traj} steps: {a.particles} "
                      f"particles, {a.steps} output steps of 0.1: kernel {st['kernel_ms']:.1f} ms, "
                      f"{st['accepted'] / (st['kernel_ms'] * 1e-3):.4e} accepted steps/s of "
                      f"kernel time, points {int(st['points'].sum())} "
                      f"({int((st['points'] == slots).sum())} particles complete), wall "
                      f"{wall:.2f}s", flush=True)
                tr.close()
                return
            t0 = time.perf_counter()
            st = tr.run_adaptive_diag(prm, dp)
            wall = time.perf_counter() - t0
            print(f"modes({a.nmodes}) dense: {a.particles} particles, {a.steps} output steps "
                  f"of 0.1: {st['accepted'] / wall:.4e} accepted steps/s, "
                  f"{st['rhs'] / wall:.4e} RHS/s, rejected {st['rejected']}, kernel "
                  f"{st['kernel_ms']:.1f} ms, wall {wall:.2f}s", flush=True)
            tr.close()
            return
        # --segments K: K runs of --tgyro gyroperiods each, every run taking
        # the resident particles on from where the last one left them (the
        # controller restarts from its first trial step in each)
        tot = {"accepted": 0, "rhs": 0, "rejected": 0, "failed": 0, "wall": 0.0}
        for seg in range(max(1, a.segments)):
            t0 = time.perf_counter()
            st = tr.run_adaptive(prm)
            wall = time.perf_counter() - t0
            for k in ("accepted", "rhs", "rejected"):
                tot[k] += st[k]
            tot["failed"] = max(tot["failed"], st["failed"])
            tot["wall"] += wall
            print(f"modes({a.nmodes}): {a.particles} particles, {a.tgyro} T_g"
                  f"{f' (segment {seg + 1} of {a.segments})' if a.segments > 1 else ''}: "
                  f"{st['accepted'] / wall:.4e} accepted steps/s, {st['rhs'] / wall:.4e} RHS/s, "
                  f"rejected {st['rejected']}, failed {st['failed']}, kernel "
                  f"{st['kernel_ms']:.1f} ms, wall {wall:.2f}s", flush=True)
        if a.segments > 1:
            ptl_end, nst = tr.store()
            g0 = np.sqrt(1.0 + np.sum(ptl[:, 4:7] ** 2, axis=1))        # ux uy uz
            g1 = np.sqrt(1.0 + np.sum(ptl_end[:, 4:7] ** 2, axis=1))
            print(f"modes({a.nmodes}): {a.particles} particles to {a.segments * a.tgyro} T_g: "
                  f"{tot['accepted'] / tot['wall']:.4e} accepted steps/s, "
                  f"{tot['rhs'] / tot['wall']:.4e} RHS/s, accepted {tot['accepted']}, "
                  f"rejected {tot['rejected']}, failed {tot['failed']}, wall "
                  f"{tot['wall']:.1f}s; max |gamma/gamma0 - 1| {np.max(np.abs(g1 / g0 - 1)):.2e}",
                  flush=True)
        tr.close()
        return
    if a.what == "sursweep":
        sur_sweep(tr, a.epochs if a.epochs > 0 else 4)
        tr.close()
        return
    if a.what == "surrogate":
        from . import surrogate
        rows, period = surrogate.surrogate_field(a.nmodes)
        tr.set_field_modes(rows, 1.0)
        q = surrogate.fourier_frequencies(period, sigma=3.0, seed=3)
        tr.mlp_set(q, surrogate.init_params(3), surrogate.output_scale(rows))
        if a.epochs > 0:
            st = tr.train_surrogate(abi.make_mlp_train_params(
                nsamples=1 << 24, period=period, batch=16384, epochs=a.epochs))
            print(f"trained: {st}", flush=True)
            # residual |B_net - B| / |B| at 65,536 uniform points
            xs = np.random.default_rng(4).uniform(0.0, period, (1 << 16, 3))
            pts = np.zeros((xs.shape[0], 4))
            pts[:, :3] = xs
            exact = tr.eval_field(pts)[:, :3]
            bnet = surrogate.output_scale(rows) * tr.mlp_forward(xs)
            bnet[:, 2] += 1.0
            res = np.linalg.norm(bnet - exact, axis=1) / np.linalg.norm(exact, axis=1)
            qs = np.quantile(res, [0.01, 0.1, 0.5, 0.9, 0.99])
            print(f"residual after {a.epochs} epochs: p1 {qs[0]:.3e} p10 {qs[1]:.3e} "
                  f"median {qs[2]:.3e} p90 {qs[3]:.3e} p99 {qs[4]:.3e}; below 1e-3: "
                  f"{np.mean(res < 1e-3):.4f}", flush=True)
        if not a.exact:
            tr.set_field_mlp(a.tol, 100)
        lo, hi = (0.0, 0.0, 0.0), (period,) * 3
        ptl = init_particles(a.particles, lo, hi, 0.7)
        tr.load(ptl)
        prm = abi.make_params(dt=0.01 * gyroperiod(0.7, 1.0), estep=a.steps, lo=lo, hi=hi,
                              pusher=abi.TP_PUSH_BORIS, precision=abi.TP_FP32_MIXED,
                              nbins=102, nt_out=max(1, a.steps // 100), emin=1e-4, emax=1.0)
        t0 = time.perf_counter()
        tr.run(prm)
        wall = time.perf_counter() - t0
        ms, k = tr.last_kernel_time()
        nfb = tr.mlp_fallback_count() if not a.exact else 0
        print(f"surrogate({'exact' if a.exact else 'net'}, tol {a.tol}): {a.particles} x "
              f"{a.steps} steps, {a.particles * a.steps / wall:.4e} particle-steps/s, "
              f"kernels {ms:.2f} ms / {k} launches, fallbacks {nfb}", flush=True)
        tr.close()
        return
    if a.what == "pmap":
        # the Poincare workload of bench.py: 256^3 turbulent grid, plane z = 128.3
        from .gpu import pmap_starts
        _, b, lo, hi = pic_turbulence_grid(256)
        tr.set_field_grid(np.zeros_like(b), b, lo, hi, abi.TP_GATHER_REF_COMPAT)
        st = pmap_starts(2, (0.0, 0.0, 128.3), min_h=lo[0] + 0.25, max_h=hi[0] - 0.25,
                         min_v=lo[1] + 0.25, max_v=hi[1] - 0.25, nh=a.side, nv=a.side)
        prm = abi.make_pmap_params(face=2, ncross=a.ncross, step=0.05, length_max=2e4, lo=lo,
                                   hi=hi, chunk=a.chunk, lanes=a.lanes)
        t0 = time.perf_counter()
        r = tr.poincare_map(prm, st)
        wall = time.perf_counter() - t0
        ms, k = tr.last_kernel_time()
        steps = int(r["nsteps"].sum())
        print(f"pmap: {st.shape[0]} lines x {a.ncross} crossings, lanes {a.lanes}, "
              f"{steps} RK4 steps, {steps / wall:.4e} steps/s, wall {wall:.3f} s, "
              f"kernels {ms:.2f} ms / {k} launches", flush=True)
        tr.close()
        return
    if a.what == "fieldload":
        field_load(tr, a.side, a.h5)
        tr.close()
        return
    pusher = abi.TP_PUSH_BORIS if a.pusher == "boris" else abi.TP_PUSH_WIRZ
    prec = abi.TP_FP32_MIXED if a.fp32 else abi.TP_FP64
    if a.what == "grid":
        tr.set_grid_cells({"auto": abi.TP_CELLS_AUTO, "off": abi.TP_CELLS_OFF,
                           "on": abi.TP_CELLS_ON}[a.cells])
        e, b, lo, hi = pic_turbulence_grid(256)
        if a.two:
            e0, b0, _, _ = pic_turbulence_grid(256, seed=20261203)
            tr.set_field_grid2(e, b, e0, b0, lo, hi, 20.0)
        else:
            tr.set_field_grid(e, b, lo, hi, abi.TP_GATHER_REF_COMPAT)
        dt = 0.01 * gyroperiod(0.7, 1.0)
    else:
        lo, hi = (0.0, 0.0, 0.0), (2 * np.pi,) * 3
        tr.set_field_abc(1.0, 1.0, 1.0)
        dt = 0.01 * gyroperiod(0.7, abc_b_ref())
    ptl = init_particles(a.particles, lo, hi, 0.7)
    if a.sort and a.what == "grid":
        # cell index i + n (j + n k) of the start point (x fastest, as the
        # nodes and cell records are laid out)
        n = 256
        c = np.clip(np.floor(ptl[:, :3] - np.asarray(lo)).astype(np.int64), 0, n - 1)
        ptl = ptl[np.argsort(c[:, 0] + n * (c[:, 1] + n * c[:, 2]), kind="stable")]
    prm = abi.make_params(dt=dt, estep=a.steps, lo=lo, hi=hi, pusher=pusher, precision=prec,
                          nbins=102, nt_out=max(1, a.steps // 100), emin=1e-4, emax=1.0,
                          dcoeff_interval=a.dinterval)
    for rep in range(a.reps):
        tr.load(ptl)
        t0 = time.perf_counter()
        tr.run(prm)
        wall = time.perf_counter() - t0
        ms, k = tr.last_kernel_time()
        print(f"{a.what} #{rep}: {a.particles} x {a.steps} steps, "
              f"{a.particles * a.steps / wall:.4e} particle-steps/s, kernels {ms:.2f} ms / "
              f"{k} launches", flush=True)
    tr.close()
    sys.stdout.flush()


if __name__ == "__main__":
    main()
