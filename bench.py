"""Benchmark of the tracking hot path on one MI355X.

Workload (the PIC-grid configuration): a 256^3 periodic turbulent E/B grid
(stored on 257^3 endpoint-inclusive nodes, FP32, 6 components), B = z-hat +
dB with a k^-5/3 spectrum and rms |dB| = 1, E = -u x B with rms |u| = 0.05c;
1,048,576 protons (q/m = 1, c = 1) at |v| = 0.7c, isotropic; FP64 Boris push
with the reference's trilinear gather; dt = 0.01 gyroperiods (T_g = 2 pi
gamma0 / B0); spectrum in 102 bins and diffusion coefficients every 100
steps, as the reference's loop computes them.

A "step" is one time step of every particle.  Inputs are resident in HBM
before timing starts; the library call is synchronous, so the wall clock
around it brackets exactly the K timed steps.

The headline value is the chosen form: the FP32 push (FP64 position and
energy accumulators) when its D_xx and its energy spectra (relative L1
distance) stay within 1 % of the FP64 run on the same particles in the same
invocation, else the plain FP64 push; the plain value is reported beside
it.  Extra workloads under "forms": the ABC field
(16,384 particles x 10,000 steps FP64 Boris and Wirz; 4M particles x 5,000
steps FP32 against FP64) and the batched adaptive Bulirsch-Stoer stepper
(1,000-mode turbulence, 1M particles), the learned-surrogate workload, the
PIC grid with two time slices and the Poincare map of the PIC grid's B
(262,144 field lines).

    python bench.py [--gpus 1] [--steps 5000] [--warmup 200]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s on 1 MI355X, chosen vs plain form, at D_xx err <=1%"
HBM_PEAK_GBS = 8000.0
BYTES_PER_PSTEP = 192.0      # 8 corners x 6 FP32 components per field evaluation
# VALU issue: one wave64 vector instruction per SIMD every 4 cycles (16 lanes
# per SIMD, the FP64 and unpacked-FP32 rate), 256 CUs x 4 SIMDs at the 2.4 GHz
# peak engine clock
VALU_PEAK_GIPS = 256 * 4 * 2.4 / 4.0
# the FP32 form is chosen only if it passes this validation run outside the
# timed region (ten D rows, so the gate is the same whatever --steps is)
VALIDATION_STEPS = 1000
COUNTERS_JSON = os.path.join(ROOT, "profiles", "k_track_grid_counters.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--particles", type=int, default=1 << 20)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 form")
    ap.add_argument("--forms", default="abc,bs,surrogate,gridwirz,grid2,openbc,poincare,fieldload",
                    help="extra workloads reported under 'forms' (abc, bs, surrogate, gridwirz, "
                         "grid2, openbc, poincare, fieldload; '' for none)")
    ap.add_argument("--pmap-side", type=int, default=512,
                    help="Poincare workload: start points per side of the plane")
    ap.add_argument("--pmap-ncross", type=int, default=8, help="Poincare crossings per line")
    ap.add_argument("--sur-epochs", type=int, default=4, help="surrogate training epochs")
    ap.add_argument("--sur-sigma", type=float, default=4.0,
                    help="spread of the surrogate's Fourier frequencies (lattice units)")
    ap.add_argument("--sur-steps", type=int, default=5000, help="surrogate-workload steps")
    ap.add_argument("--bs-tgyro", type=float, default=2.0,
                    help="adaptive-stepper horizon in gyroperiods")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {args.gpus}); the "
              "particles would shard across GPUs without communication during the run",
              file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: one process per run; WORLD_SIZE > 1 is not supported", file=sys.stderr)
        sys.exit(2)

    from test_particle_amd import abi
    from test_particle_amd.fields import gyroperiod, pic_turbulence_grid
    from test_particle_amd.gpu import Tracker
    from test_particle_amd.particles import init_particles

    t0 = time.perf_counter()
    log(f"[bench] building {args.grid}^3 turbulence grid (seed {args.seed})")
    e, b, lo, hi = pic_turbulence_grid(args.grid, seed=args.seed)
    log(f"[bench] grid ready in {time.perf_counter() - t0:.1f}s")
    n = args.particles
    vthe = 0.7
    dt = 0.01 * gyroperiod(vthe, 1.0)
    ptl = init_particles(n, lo, hi, vthe, seed=args.seed)

    def params(steps, precision=abi.TP_FP64):
        return abi.make_params(dt=dt, estep=steps, lo=lo, hi=hi, pusher=abi.TP_PUSH_BORIS,
                               precision=precision, nbins=102,
                               nt_out=max(1, min(100, steps // 100)), emin=1e-4, emax=1.0)

    tr = Tracker(0)
    gpu_name, ncu = tr.device_info()
    log(f"[bench] {gpu_name}, {ncu} CUs")
    tr.set_field_grid(e, b, lo, hi, abi.TP_GATHER_REF_COMPAT)
    gate = None
    if not args.no_fp32:
        gate = validate_fp32(tr, ptl, params, VALIDATION_STEPS)
    tr.load(ptl)
    if args.warmup > 0:
        log(f"[bench] warmup {args.warmup} steps")
        tr.run(params(args.warmup))
        tr.load(ptl)                  # the timed run starts from the initial particles
    K = args.steps
    prm = params(K)
    log(f"[bench] timed {K} steps x {n} particles (FP64 Boris, reference gather)")
    t1 = time.perf_counter()
    tr.run(prm)                       # synchronous: returns after the last kernel
    wall = time.perf_counter() - t1
    kms, launches = tr.last_kernel_time()
    value = n * K / wall
    diag = tr.diagnostics(prm)
    dxx = diag["dcoeff"][1, 1:] / np.maximum(diag["remain"][1:], 1)
    log(f"[bench] {value:.4e} particle-steps/s ({wall:.3f}s wall, kernels {kms:.1f} ms "
        f"in {launches} launches)")

    forms = {"plain_fp64": {"value": value, "dtype": "f64", "kernel_ms": kms,
                            "launches": launches}}
    chosen, chosen_value, chosen_dtype = "plain_fp64", value, "f64"
    if not args.no_fp32:
        tr.load(ptl)
        prm32 = params(K, abi.TP_FP32_MIXED)
        t2 = time.perf_counter()
        tr.run(prm32)
        wall32 = time.perf_counter() - t2
        kms32, launches32 = tr.last_kernel_time()
        d32 = tr.diagnostics(prm32)
        dxx32 = d32["dcoeff"][1, 1:] / np.maximum(d32["remain"][1:], 1)
        # the timed runs' own comparison (no D rows when --steps is below one
        # D interval); the choice rests on the validation run
        err = float(np.max(np.abs(dxx32 / dxx - 1.0))) if dxx.size else None
        serr = spectrum_err(d32["espect"], diag["espect"])
        v32 = n * K / wall32
        forms["fp32_mixed"] = {"value": v32, "dtype": "f32 push and gather, f64 position "
                                                     "and energy accumulators",
                               "validation": gate,
                               "timed_run_dxx_max_rel_err_vs_fp64": err,
                               "timed_run_spectrum_l1_rel_err_vs_fp64": serr,
                               "kernel_ms": kms32, "launches": launches32}
        log(f"[bench] FP32 form {v32:.4e} particle-steps/s; validation over "
            f"{gate['steps']} steps: D_xx err {gate['dxx_max_rel_err']:.2e}, spectrum L1 err "
            f"{gate['spectrum_l1_rel_err']:.2e} -> {'passes' if gate['passes'] else 'fails'}")
        if gate["passes"] and v32 > value:
            chosen, chosen_value = "fp32_mixed", v32
            chosen_dtype = "f32 (f64 position/energy accumulators)"
            kms, launches = kms32, launches32
    roofline = headline_roofline(chosen, kms, launches, n, K)
    # what this GPU reaches on a plain 16-byte-per-lane HBM copy (2 GiB each
    # way), beside the published peak
    from test_particle_amd.gpu import measure_peak
    copy_gbs, _ = measure_peak(abi.TP_PEAK_HBM_COPY, 2 << 30, 3)
    roofline["hbm"]["peak_measured_copy"] = copy_gbs
    log(f"[bench] HBM copy microbenchmark {copy_gbs:.0f} GB/s; roofline bound "
        f"{roofline['bound']}, frac {roofline['frac']}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(e, b, lo, hi, ptl, params, tr)
    tr.close()
    del e, b
    extra = [f.strip() for f in args.forms.split(",") if f.strip()]
    if "abc" in extra:
        forms.update(forms_abc(args))
    if "bs" in extra:
        forms.update(forms_bs(args))
    if "surrogate" in extra:
        forms.update(forms_surrogate(args))
    if "fieldload" in extra:
        forms.update(forms_fieldload(args, not args.no_cpu))
    if "gridwirz" in extra:
        forms.update(forms_gridwirz(args, not args.no_cpu))
    if "grid2" in extra:
        forms.update(forms_grid2(args))
    if "openbc" in extra:
        forms.update(forms_openbc(args))
    if "poincare" in extra:
        forms.update(forms_poincare(args, not args.no_cpu))

    out = {
        "metric": METRIC, "value": chosen_value, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": K, "warmup": args.warmup, "ms_per_step": n * 1e3 / chosen_value,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": chosen_dtype, "data": "synthetic",
        "config": {"workload": f"PIC-grid fields: {args.grid}^3 periodic turbulent FP32 E,B "
                               f"({args.grid + 1}^3 endpoint-inclusive nodes), {n} protons "
                               f"|v|=0.7c, Boris, reference trilinear gather, dt=0.01 T_g, "
                               f"{K} steps",
                   "particles": n, "grid_nodes": [args.grid + 1] * 3, "pusher": "boris",
                   "gather": "ref_compat", "seed": args.seed, "gpu": gpu_name,
                   "chosen_form": chosen, "fp32_gate": gate, "plain_value": value,
                   "chosen_vs_plain": chosen_value / value,
                   "dxx_last": float(dxx[-1]) if dxx.size else None},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(out), flush=True)


def validate_fp32(tr, ptl, params, steps):
    """The FP32 form against the FP64 form on the same particles over
    `steps` steps (outside the timed region): largest relative D_xx error
    over the D rows and largest relative L1 distance of the spectra."""
    from test_particle_amd import abi
    out = {}
    for name, prec in (("f64", abi.TP_FP64), ("f32", abi.TP_FP32_MIXED)):
        prm = params(steps, prec)
        tr.load(ptl)
        tr.run(prm)
        d = tr.diagnostics(prm)
        out[name] = (d["dcoeff"][1, 1:] / np.maximum(d["remain"][1:], 1), d["espect"])
    err = float(np.max(np.abs(out["f32"][0] / out["f64"][0] - 1.0)))
    serr = spectrum_err(out["f32"][1], out["f64"][1])
    return {"steps": steps, "dxx_rows": int(out["f64"][0].size), "dxx_max_rel_err": err,
            "spectrum_l1_rel_err": serr, "passes": bool(err <= 0.01 and serr <= 0.01)}


def headline_roofline(form, kms, launches, n, K):
    """Roofline of the chosen form's step kernel against the resource that
    binds it.  Per launch, its VALU instructions and HBM bytes come from the
    rocprofv3 counters in profiles/k_track_grid_counters.json, measured at
    launch lengths of 20 and 100 steps (fixed part + per-step part; this
    run's average launch length picks the point on that line); its duration
    is the live HIP-event time of this run's launches.  VALU: instructions
    issued per second against one wave64 instruction per SIMD per 4 cycles;
    HBM: counter bytes per second against 8 TB/s.  The bound is the larger
    fraction; the algorithmic bytes (192 per particle-step, no reuse credited)
    are reported beside, with their ratio to the counted traffic."""
    spl = K / max(launches, 1)
    t_launch = kms / max(launches, 1) * 1e-3            # seconds
    alg = BYTES_PER_PSTEP * n * spl                      # algorithmic bytes per launch
    r = {"bound": None, "achieved": None, "peak": None, "unit": None, "frac": None,
         "traffic": None,
         "kernel": "k_track<BORIS, GRID_REF, %s>" % ("FP64" if form == "plain_fp64" else "FP32"),
         "kernel_ms_per_launch": t_launch * 1e3, "steps_per_launch": spl,
         "valu": None, "hbm": {"peak": HBM_PEAK_GBS, "unit": "GB/s"},
         "algorithmic": {"bytes_per_particle_step": BYTES_PER_PSTEP,
                         "GBps": alg / t_launch / 1e9 if t_launch > 0 else None}}
    if not os.path.exists(COUNTERS_JSON) or t_launch <= 0:
        return r
    with open(COUNTERS_JSON) as f:
        cj = json.load(f)
    ent = cj.get(form)
    if not ent or ent.get("particles") != n:
        return r
    pts = sorted((int(k), v) for k, v in ent["per_launch"].items())

    def at(key):
        (s0, a), (s1, b) = pts[0], pts[-1]
        if s1 == s0:
            return a[key]
        slope = (b[key] - a[key]) / (s1 - s0)
        return a[key] + slope * (spl - s0)

    valu = at("valu_insts")
    traffic = at("hbm_bytes")
    valu_rate = valu / t_launch / 1e9
    hbm_rate = traffic / t_launch / 1e9
    fv, fh = valu_rate / VALU_PEAK_GIPS, hbm_rate / HBM_PEAK_GBS
    r["valu"] = {"achieved": valu_rate, "peak": VALU_PEAK_GIPS,
                 "unit": "G wave64 instructions/s", "frac": fv,
                 "insts_per_launch": valu, "insts_per_wave_step": valu / (n / 64.0) / spl}
    r["hbm"].update({"achieved": hbm_rate, "frac": fh, "bytes_per_launch": traffic})
    r["algorithmic"]["ratio_traffic_to_algorithmic"] = traffic / alg
    r["traffic"] = traffic
    r["source"] = os.path.relpath(COUNTERS_JSON, ROOT)
    if fv >= fh:
        r.update(bound="valu", achieved=valu_rate, peak=VALU_PEAK_GIPS,
                 unit="G wave64 instructions/s", frac=fv)
    else:
        r.update(bound="hbm", achieved=hbm_rate, peak=HBM_PEAK_GBS, unit="GB/s", frac=fh)
    return r


def spectrum_err(a, b):
    """Largest relative L1 distance between two energy spectra over the
    frames: sum_bins |a - b| / sum_bins b (0 for identical spectra; the
    fraction of particles binned differently, counted twice)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    tot = b.sum(axis=1)
    m = tot > 0
    if not m.any():
        return 0.0
    return float(np.max(np.abs(a - b).sum(axis=1)[m] / tot[m]))


def forms_abc(args):
    """The ABC force-free field workloads: 16,384 particles x 10,000 steps
    (FP64 Boris and FP64 Wirz), and 4M particles x 5,000 steps with the FP32
    push against the FP64 push on identical inputs (per-particle energy
    drift, D_xx and spectrum compared)."""
    from test_particle_amd import abi
    from test_particle_amd.fields import abc_b_ref, gyroperiod
    from test_particle_amd.gpu import Tracker
    from test_particle_amd.particles import init_particles
    out = {}
    lo, hi = (0.0, 0.0, 0.0), (2 * np.pi,) * 3
    vthe = 0.7
    dt = 0.01 * gyroperiod(vthe, abc_b_ref())
    g0 = 1.0 / np.sqrt(1.0 - vthe * vthe)

    def params(steps, pusher, precision):
        return abi.make_params(dt=dt, estep=steps, lo=lo, hi=hi, pusher=pusher,
                               precision=precision, nbins=102, nt_out=100, emin=1e-3,
                               emax=1e1)

    with Tracker(0) as tr:
        tr.set_field_abc(1.0, 1.0, 1.0)
        n1, k1 = 16384, 10000
        p1 = init_particles(n1, lo, hi, vthe, seed=args.seed)
        for name, push in (("abc16k_fp64_boris", abi.TP_PUSH_BORIS),
                           ("abc16k_fp64_wirz", abi.TP_PUSH_WIRZ)):
            tr.load(p1)
            prm = params(k1, push, abi.TP_FP64)
            t0 = time.perf_counter()
            tr.run(prm)
            w = time.perf_counter() - t0
            d = tr.diagnostics(prm)
            out[name] = {"value": n1 * k1 / w, "dtype": "f64", "particles": n1, "steps": k1,
                         "dxx_last": float(d["dcoeff"][1, -1] / max(d["remain"][-1], 1))}
            log(f"[bench] {name}: {n1 * k1 / w:.4e} particle-steps/s")
            if push == abi.TP_PUSH_WIRZ and not args.no_cpu:
                base = abc_reference_baseline(p1, dt, k1, d)
                if base is not None:
                    out[name]["cpu_baseline"] = base
                    log(f"[bench] {name} CPU (reference) {base['value']:.3e} particle-steps/s "
                        f"on {base['cores']} threads; D_xx last row GPU vs reference "
                        f"{base['dxx_last_rel_diff_vs_gpu']:.2e}")
        n3, k3 = 1 << 22, 5000
        p3 = init_particles(n3, lo, hi, vthe, seed=args.seed)
        res = {}
        for prec in (abi.TP_FP64, abi.TP_FP32_MIXED):
            tr.load(p3)
            prm = params(k3, abi.TP_PUSH_BORIS, prec)
            t0 = time.perf_counter()
            tr.run(prm)
            w = time.perf_counter() - t0
            fin, _ = tr.store()
            res[prec] = (n3 * k3 / w, tr.diagnostics(prm), fin)
        v64, d64, f64 = res[abi.TP_FP64]
        v32, d32, f32 = res[abi.TP_FP32_MIXED]
        dxx64 = d64["dcoeff"][1, 1:] / np.maximum(d64["remain"][1:], 1)
        dxx32 = d32["dcoeff"][1, 1:] / np.maximum(d32["remain"][1:], 1)

        def gam(p):
            return np.sqrt(1.0 + p[:, 4] ** 2 + p[:, 5] ** 2 + p[:, 6] ** 2)
        drift32 = np.abs(gam(f32) / g0 - 1.0)
        drift64 = np.abs(gam(f64) / g0 - 1.0)
        out["abc4M_fp64_boris"] = {"value": v64, "dtype": "f64", "particles": n3, "steps": k3,
                                   "energy_drift_max": float(drift64.max())}
        out["abc4M_fp32_boris"] = {
            "value": v32, "dtype": "f32 push and field, f64 position/energy accumulators",
            "particles": n3, "steps": k3, "speedup_vs_fp64": v32 / v64,
            "dxx_max_rel_err_vs_fp64": float(np.max(np.abs(dxx32 / dxx64 - 1.0))),
            "spectrum_l1_rel_err_vs_fp64": spectrum_err(d32["espect"], d64["espect"]),
            "energy_drift_max": float(drift32.max()),
            "energy_drift_median": float(np.median(drift32))}
        log(f"[bench] ABC 4M: FP64 {v64:.4e}, FP32 {v32:.4e} particle-steps/s, "
            f"D_xx err {out['abc4M_fp32_boris']['dxx_max_rel_err_vs_fp64']:.2e}, "
            f"FP32 energy drift max {drift32.max():.2e}")
    return out


def forms_bs(args):
    """Batched adaptive Bulirsch-Stoer: 1000-mode turbulence (k over 3
    decades, |dB_k| ~ k^-5/3, rms 0.5 over B0 = 1), 1M particles with |v|
    log-uniform in [0.1, 0.9]c, rtol 1e-8, each particle integrated over
    `--bs-tgyro` of its own gyroperiods in one persistent work-queue
    launch.  One accepted BS step is counted as one particle-step."""
    from test_particle_amd import abi
    from test_particle_amd.fields import turbulence_modes
    from test_particle_amd.gpu import Tracker
    from test_particle_amd.particles import init_particles
    n = 1 << 20
    rows = turbulence_modes(1000, seed=args.seed)
    ptl = init_particles(n, (0, 0, 0), (100, 100, 100), 0.1, kind="loguniform", vmax=0.9,
                         seed=args.seed)
    prm = abi.make_bs_params(rtol=1e-8, atol=1e-10, t_end_gyro=args.bs_tgyro, b_ref=1.0)
    with Tracker(0) as tr:
        tr.set_field_modes(rows, 1.0)
        tr.load(ptl[:65536])
        tr.run_adaptive(abi.make_bs_params(rtol=1e-8, atol=1e-10, t_end_gyro=0.01,
                                           b_ref=1.0))
        tr.load(ptl)
        t0 = time.perf_counter()
        st = tr.run_adaptive(prm)
        w = time.perf_counter() - t0
        fin = tr.store()[0] if not args.no_cpu else None
        # the rate over horizons of 0.2, 2 and 20 gyroperiods, all particles
        # (the persistent kernel's lanes are refilled from the same queue)
        flat, flat_rhs = {}, {}
        for tg in (0.2, 2.0, 20.0):
            if tg == args.bs_tgyro:
                flat[str(tg)], flat_rhs[str(tg)] = st["accepted"] / w, st["rhs"] / w
                continue
            tr.load(ptl)
            t1 = time.perf_counter()
            s2 = tr.run_adaptive(abi.make_bs_params(rtol=1e-8, atol=1e-10, t_end_gyro=tg,
                                                    b_ref=1.0))
            w2 = time.perf_counter() - t1
            flat[str(tg)], flat_rhs[str(tg)] = s2["accepted"] / w2, s2["rhs"] / w2
        log(f"[bench] BS accepted steps/s over 0.2/2/20 T_g: {flat}; RHS/s {flat_rhs}")
        dense = bs_dense_vs_fixed(tr, abi, ptl)
    rate = st["accepted"] / w
    log(f"[bench] BS: {rate:.4e} accepted steps/s, {st['rhs'] / w:.4e} RHS/s, "
        f"rejected {st['rejected']}, failed {st['failed']}")
    out = {"bs_modes1000_1M": {
        "value": rate, "unit": "accepted BS steps/s", "dtype": "f64", "particles": n,
        "horizon_gyroperiods": args.bs_tgyro, "rtol": 1e-8, "atol": 1e-10,
        "rhs_per_s": st["rhs"] / w, "accepted": st["accepted"], "rejected": st["rejected"],
        "failed": st["failed"], "kernel_ms": st["kernel_ms"], "wall_s": w,
        "rate_by_horizon_gyroperiods": flat, "rhs_rate_by_horizon_gyroperiods": flat_rhs}}
    out.update(dense)
    if fin is not None:
        # the oracle's sequential BS (OpenMP over particles) on the first
        # particles of the same set, about 10 s; the GPU's end states of
        # those particles are compared bit for bit
        from oracle import tpref
        cores = host_cores()
        tpref.set_math(False)
        f = tpref.Field.modes(rows, 1.0)
        pilot = tpref.bs_track(f, prm, ptl[:8 * cores], nthreads=cores)
        per = max(pilot["seconds"], 1e-6) / (8 * cores)          # thread-seconds / particle
        m = int(np.clip(10.0 * cores / (per * cores), 8 * cores, n))
        ref = tpref.bs_track(f, prm, ptl[:m], nthreads=cores)
        same = bool(np.array_equal(ref["ptl"], fin[:m]))
        out["bs_modes1000_1M"]["cpu_baseline"] = {
            "value": ref["accepted"] / ref["seconds"], "unit": "accepted BS steps/s",
            "cores": cores, "kind": "port",
            "sample": f"first {m} particles of the same set, same horizon and tolerances, "
                      f"the oracle's BS with OpenMP ({cpu_model()})",
            "gpu_sample_bitwise_equal": same}
        log(f"[bench] BS CPU {ref['accepted'] / ref['seconds']:.3e} accepted steps/s on "
            f"{cores} threads ({m} particles); GPU sample bitwise={same}")
    return out


def bs_dense_vs_fixed(tr, abi, ptl):
    """The adaptive stepper with dense output and the fixed-step loop's
    diagnostics (tp_gpu_run_adaptive_diag) on the same 1000-mode field and
    1M particles, against the plain FP64 Boris kernel on the same output
    grid: dt = 0.1 (0.007-0.016 gyroperiods), 200 steps (t = 20, 1.4-3.2
    gyroperiods), 10 spectrum frames, D rows every 20 steps, no boundary.
    D_xx is compared row by row, with the standard error of the last row's
    difference from the per-particle terms; spectra by relative L1."""
    n = ptl.shape[0]
    dt, estep = 0.1, 200
    prm = abi.make_params(dt=dt, estep=estep, lo=(0, 0, 0), hi=(100, 100, 100),
                          pusher=abi.TP_PUSH_BORIS, bc_flag=abi.TP_BC_NONE, nbins=102,
                          nt_out=10, emin=1e-4, emax=1.0, dcoeff_interval=20)
    tr.load(ptl)
    t0 = time.perf_counter()
    tr.run(prm)
    wf = time.perf_counter() - t0
    pf, _ = tr.store()
    df = tr.diagnostics(prm)
    bs = abi.make_bs_params(rtol=1e-8, atol=1e-10, b_ref=1.0)
    tr.load(ptl)
    t0 = time.perf_counter()
    st = tr.run_adaptive_diag(bs, prm)
    wa = time.perf_counter() - t0
    pa, _ = tr.store()
    da = tr.diagnostics(prm)
    dxx = lambda d: d["dcoeff"][1, 1:] / np.maximum(d["remain"][1:], 1)  # noqa: E731
    rel = np.abs(dxx(da) / dxx(df) - 1.0)
    t_end = estep * dt

    def terms(p):
        return (ptl[:, 0] - (p[:, 0] + p[:, 8])) ** 2 / (2 * t_end + dt)
    a, b = terms(pf), terms(pa)
    se = float(np.hypot(a.std(ddof=1), b.std(ddof=1)) / np.sqrt(n))
    serr = spectrum_err(da["espect"], df["espect"])
    log(f"[bench] BS dense + diagnostics: {st['accepted'] / wa:.4e} accepted steps/s; D_xx vs "
        f"fixed-step FP64 max rel {rel.max():.2e} (last row {rel[-1]:.2e}, "
        f"{abs(b.mean() - a.mean()) / se:.2f} SE); spectrum L1 {serr:.2e}")
    return {"bs_dense_diag_modes1000_1M": {
        "value": st["accepted"] / wa, "unit": "accepted BS steps/s", "dtype": "f64",
        "particles": n, "output_dt": dt, "output_steps": estep, "rtol": 1e-8, "atol": 1e-10,
        "rhs_per_s": st["rhs"] / wa, "accepted": st["accepted"], "rejected": st["rejected"],
        "failed": st["failed"], "kernel_ms": st["kernel_ms"],
        "vs_fixed_step_fp64": {"fixed_step_value": n * estep / wf,
                               "dxx_max_rel_diff": float(rel.max()),
                               "dxx_last_rel_diff": float(rel[-1]),
                               "dxx_last_diff_in_se": float(abs(b.mean() - a.mean()) / se),
                               "dxx_last_se_rel": se / float(a.mean()),
                               "spectrum_l1_rel_diff": serr}}}


def forms_surrogate(args):
    """The learned-surrogate workload: the 1000-mode turbulence of the
    adaptive workload with its wavevectors on the lattice of its own k_min
    (periodic box L = 2 pi / k_min), the network trained on the GPU
    (2^24 exact FP64 samples), then 4M particles at |v| = 0.7c, FP32 Boris,
    dt = 0.01 T_g, periodic in L: the exact FP32 mode sum against the
    network with the exact-field fallback at tolerance 1e-3 (checked every
    100 steps) and, as the plain form, the FP64 push with the exact FP64 mode
    sum (timed over 500 steps).  D_xx and the spectrum are compared with
    the exact-field FP32 run."""
    from test_particle_amd import abi, surrogate
    from test_particle_amd.fields import gyroperiod
    from test_particle_amd.gpu import Tracker
    from test_particle_amd.particles import init_particles
    rows, period = surrogate.surrogate_field(1000, seed=args.seed)
    lo, hi = (0.0, 0.0, 0.0), (period,) * 3
    n, K = 1 << 22, args.sur_steps
    ptl = init_particles(n, lo, hi, 0.7, seed=args.seed)
    dt = 0.01 * gyroperiod(0.7, 1.0)

    def params(steps, prec):
        return abi.make_params(dt=dt, estep=steps, lo=lo, hi=hi, pusher=abi.TP_PUSH_BORIS,
                               precision=prec, nbins=102, nt_out=max(1, min(100, steps // 100)),
                               emin=1e-4, emax=1.0)
    out = {}
    with Tracker(0) as tr:
        tr.set_field_modes(rows, 1.0)
        q = surrogate.fourier_frequencies(period, sigma=args.sur_sigma, seed=args.seed)
        scale = surrogate.output_scale(rows)
        tr.mlp_set(q, surrogate.init_params(args.seed), scale)
        tp = abi.make_mlp_train_params(nsamples=1 << 24, period=period, batch=16384,
                                       epochs=args.sur_epochs, lr=2e-3, lr_min=2e-5,
                                       seed=args.seed)
        log(f"[bench] surrogate: training {args.sur_epochs} epochs of 2^24 samples")
        st = tr.train_surrogate(tp)
        log(f"[bench] surrogate trained: loss {st['first_loss']:.3e} -> {st['last_loss']:.3e}, "
            f"{st['steps']} steps in {st['train_ms'] / 1e3:.1f}s (samples {st['sample_ms']:.0f} ms)")
        xs = np.random.default_rng(args.seed + 1).uniform(0.0, period, (1 << 16, 3))
        pts = np.zeros((xs.shape[0], 4))
        pts[:, :3] = xs
        exact = tr.eval_field(pts)[:, :3]
        bnet = scale * tr.mlp_forward(xs)
        bnet[:, 2] += 1.0
        res = np.linalg.norm(bnet - exact, axis=1) / np.linalg.norm(exact, axis=1)
        # plain FP64 form (exact FP64 mode sum), a shorter run
        k64 = min(K, 500)
        tr.load(ptl)
        t0 = time.perf_counter()
        tr.run(params(k64, abi.TP_FP64))
        v64 = n * k64 / (time.perf_counter() - t0)
        # exact FP32 (what the surrogate is compared with)
        tr.load(ptl)
        pe = params(K, abi.TP_FP32_MIXED)
        t0 = time.perf_counter()
        tr.run(pe)
        vex = n * K / (time.perf_counter() - t0)
        dex = tr.diagnostics(pe)
        dxx_ex = dex["dcoeff"][1, 1:] / np.maximum(dex["remain"][1:], 1)
        runs = {}
        for tol in (1e-3, float(np.quantile(res, 0.99))):
            tr.set_field_mlp(tol, 100)
            tr.load(ptl)
            t0 = time.perf_counter()
            tr.run(pe)
            v = n * K / (time.perf_counter() - t0)
            d = tr.diagnostics(pe)
            dxx = d["dcoeff"][1, 1:] / np.maximum(d["remain"][1:], 1)
            runs[tol] = {"value": v, "fallback_fraction": tr.mlp_fallback_count() / n,
                         "dxx_max_rel_err_vs_exact_fp32": float(np.max(np.abs(dxx / dxx_ex - 1))),
                         "spectrum_l1_rel_err_vs_exact_fp32": spectrum_err(d["espect"],
                                                                            dex["espect"])}
            tr.set_field_modes(rows, 1.0)
            log(f"[bench] surrogate tol {tol:.2e}: {v:.4e} particle-steps/s, fallback "
                f"{runs[tol]['fallback_fraction']:.3f}, D_xx err "
                f"{runs[tol]['dxx_max_rel_err_vs_exact_fp32']:.2e}")
    log(f"[bench] surrogate workload: FP64 {v64:.4e}, exact FP32 {vex:.4e} particle-steps/s")
    tols = sorted(runs)
    out["surrogate_modes1000_4M_plain_fp64"] = {
        "value": v64, "dtype": "f64", "particles": n, "steps": k64,
        "field": "exact FP64 mode sum (1000 modes)"}
    out["surrogate_modes1000_4M_exact_fp32"] = {
        "value": vex, "dtype": "f32 push and mode sum, f64 positions", "particles": n,
        "steps": K, "field": "exact FP32 mode sum (hardware cosine)"}
    for tol in tols:
        r = runs[tol]
        r.update({"dtype": "bf16 network (f32 accumulation), f32 push, f64 positions",
                  "particles": n, "steps": K, "tolerance": tol, "check_interval": 100,
                  "speedup_vs_exact_fp32": r["value"] / vex,
                  "speedup_vs_plain_fp64": r["value"] / v64})
        out[f"surrogate_modes1000_4M_net_tol{tol:.1e}"] = r
    out["surrogate_training"] = {
        "samples": 1 << 24, "batch": 16384, "epochs": args.sur_epochs,
        "steps": st["steps"], "first_loss": st["first_loss"], "last_loss": st["last_loss"],
        "train_s": st["train_ms"] / 1e3, "sample_s": st["sample_ms"] / 1e3,
        "fourier_sigma_lattice_units": args.sur_sigma, "period": period,
        "residual_median": float(np.median(res)), "residual_p90": float(np.quantile(res, 0.9)),
        "residual_p99": float(np.quantile(res, 0.99)),
        "fraction_residual_below_1e-3": float(np.mean(res < 1e-3))}
    return out


def forms_gridwirz(args, with_cpu):
    """The headline grid with the reference's own pusher: FP64 Wirz through
    the reference gather, 1M particles x 1000 steps.  Its CPU baseline is the
    reference's particle_tracking_hybrid compiled from its sources
    (oracle/_ref/ref_harness, mode hybrid, OpenMP over the box's threads) on
    the first 65,536 particles, the step count sized from a pilot to about
    10 s; the GPU runs the same sample, and the D_xx of their last rows is
    compared statistically (the Wirz rotation's sin/cos: libm in the
    reference, the project's routine on the GPU, differ by ulps, which the
    turbulent field amplifies)."""
    from test_particle_amd import abi
    from test_particle_amd.fields import gyroperiod, pic_turbulence_grid
    from test_particle_amd.gpu import Tracker
    from test_particle_amd.particles import init_particles
    e, b, lo, hi = pic_turbulence_grid(args.grid, seed=args.seed)
    n, K = args.particles, 1000
    ptl = init_particles(n, lo, hi, 0.7, seed=args.seed)
    dt = 0.01 * gyroperiod(0.7, 1.0)

    def params(steps, nt_out):
        return abi.make_params(dt=dt, estep=steps, lo=lo, hi=hi, pusher=abi.TP_PUSH_WIRZ,
                               nbins=102, nt_out=nt_out, emin=1e-4, emax=1.0)

    out = {}
    with Tracker(0) as tr:
        tr.set_field_grid(e, b, lo, hi)
        tr.load(ptl)
        tr.run(params(100, 1))
        tr.load(ptl)
        t0 = time.perf_counter()
        tr.run(params(K, 10))
        w = time.perf_counter() - t0
        kms, launches = tr.last_kernel_time()
        name = "pic_grid_1M_fp64_wirz"
        out[name] = {"value": n * K / w, "dtype": "f64", "particles": n, "steps": K,
                     "pusher": "wirz (the reference's tracking_wirz)",
                     "kernel_ms_per_launch": kms / max(launches, 1)}
        log(f"[bench] {name}: {n * K / w:.4e} particle-steps/s")
        if with_cpu:
            ref = grid_reference_baseline(e, b, lo, hi, ptl, dt, tr, params)
            if ref is not None:
                out[name]["cpu_baseline"] = ref
                log(f"[bench] {name} CPU (reference) {ref['value']:.3e} particle-steps/s on "
                    f"{ref['cores']} threads; D_xx last row GPU vs reference "
                    f"{ref['dxx_last_rel_diff_vs_gpu']:.2e}")
    del e, b
    return out


def grid_reference_baseline(e, b, lo, hi, ptl, dt, tr, params):
    """The reference's particle_tracking_hybrid on the PIC grid (system 4,
    fixed-step Wirz) over the first particles; see forms_gridwirz.  None when
    the harness is absent."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(exe):
        return None
    cores = host_cores()
    nz, ny, nx = e.shape[:3]
    with tempfile.TemporaryDirectory() as tmp:
        os.makedirs(os.path.join(tmp, "data"))
        grid = os.path.join(tmp, "grid.bin")
        with open(grid, "wb") as f:
            np.ascontiguousarray(e, np.float32).tofile(f)
            np.ascontiguousarray(b, np.float32).tofile(f)

        def run(np_, steps):
            with open(os.path.join(tmp, "spec.txt"), "w") as f:
                f.write(f"system 4\nnx {nx}\nny {ny}\nnz {nz}\ngrid {grid}\n"
                        f"tmax {float(steps * dt + 1e-9)!r}\ndt {float(dt)!r}\n"
                        f"bc 0\nnbins 102\nnt_out {steps // 100}\nemin 0.0001\nemax 1.0\n")
                for k, a in enumerate(("x", "y", "z")):
                    f.write(f"{a}min {float(lo[k])!r}\n{a}max {float(hi[k])!r}\n")
            np.ascontiguousarray(ptl[:np_], np.float64).tofile(os.path.join(tmp, "p.bin"))
            p = subprocess.run([exe, "hybrid", "spec.txt", "p.bin"], cwd=tmp, check=True,
                               capture_output=True, text=True, timeout=600,
                               env=dict(os.environ, OMP_NUM_THREADS=str(cores)))
            kv = dict(line.split() for line in p.stdout.splitlines()
                      if line.startswith("hybrid_"))
            return float(kv["hybrid_seconds"])

        ns = 65536
        pilot = run(8192, 100)
        rate = 8192 * 100 / max(pilot, 1e-6)
        ks = int(min(max(100, round(10.0 * rate / ns / 100) * 100), 5000))
        sec = run(ns, ks)
        diff = np.loadtxt(os.path.join(tmp, "data", "diff_coeffs.dat"), skiprows=1)
    g = tr.track_fixed(params(ks, ks // 100), ptl[:ns])
    dxx_ref = diff[-1, 1] / diff[-1, 8]
    dxx_gpu = g["dcoeff"][1, -1] / g["remain"][-1]
    return {"value": ns * ks / sec, "unit": "particle-steps/s", "cores": cores,
            "kind": "reference",
            "sample": f"first {ns} of the same particles, {ks} Wirz steps, same grid, through "
                      f"the reference's particle_tracking_hybrid (oracle/_ref/ref_harness, "
                      f"OpenMP) on {cpu_model()}",
            "dxx_last_rel_diff_vs_gpu": float(abs(dxx_gpu / dxx_ref - 1.0))}


def forms_fieldload(args, with_cpu):
    """The PIC field input pipeline (tp_gpu_load_field_grid): raw component
    planes of a 384^3 grid (1.36 GB, random values) to the nodes, through
    pinned slabs and the GPU layout kernels, in the HDF5 (x-slowest) and .gda
    (x-fastest) orders, the reference's float3 arrays through
    tp_gpu_set_field_grid, and a PIC HDF5 file of the same planes
    (fieldload_h5); wall time of the second call of each (the first
    allocates the context's staging slabs), in GB/s of input."""
    from test_particle_amd import abi
    from test_particle_amd.gpu import Tracker
    n = 384
    seed = args.seed
    with_h5 = os.path.exists(abi.H5_LIB_PATH)
    rng = np.random.default_rng(seed)
    planes = rng.standard_normal((6, n, n, n), dtype=np.float32)
    lo, hi = (0.0, 0.0, 0.0), (float(n),) * 3
    gb = planes.nbytes / 1e9
    out = {}
    with Tracker(0) as tr:
        tr.set_grid_cells(abi.TP_CELLS_OFF)      # the input path alone
        for order, name in ((abi.TP_PLANES_X_SLOWEST, "fieldload_384_planes_xslowest"),
                            (abi.TP_PLANES_X_FASTEST, "fieldload_384_planes_xfastest")):
            for rep in range(2):
                t0 = time.perf_counter()
                tr.set_field_grid_planes(planes, lo, hi, order=order, e0=0.3, b0=1.7)
                w = time.perf_counter() - t0
            out[name] = {"value": gb / w, "unit": "GB/s of raw planes", "ms": w * 1e3,
                         "nodes": n ** 3}
            log(f"[bench] {name}: {gb / w:.2f} GB/s ({w * 1e3:.1f} ms)")
        e = np.ascontiguousarray(planes[:3].transpose(3, 2, 1, 0))
        b = np.ascontiguousarray(planes[3:].transpose(3, 2, 1, 0))
        del planes
        for rep in range(2):
            t0 = time.perf_counter()
            tr.set_field_grid(e, b, lo, hi)
            w = time.perf_counter() - t0
        out["fieldload_384_float3_arrays"] = {"value": gb / w, "unit": "GB/s of float3 arrays",
                                              "ms": w * 1e3, "nodes": n ** 3}
        log(f"[bench] fieldload_384_float3_arrays: {gb / w:.2f} GB/s ({w * 1e3:.1f} ms)")
        del e, b
        if with_h5:
            # the same planes as a PIC HDF5 file: onto the GPU through
            # tp_h5_load_pic, and through the reference's own loaders
            out.update(fieldload_h5(tr, n, seed, with_cpu))
        tr.set_grid_cells(abi.TP_CELLS_AUTO)
    return out


def fieldload_h5(tr, n, seed, with_cpu):
    """A 384^3 PIC HDF5 file (datasets ex ey ez cbx cby cbz, x slowest) loaded
    onto the GPU by tp_h5_load_pic (second call timed, file in the page
    cache); its CPU baseline is the reference's read_pic_efields_h5 and
    read_pic_bfields_h5 (compiled from its sources into
    oracle/_ref/ref_harness, mode h5: hyperslab reads and the transposing
    loops on one host thread), timed around the two calls."""
    import subprocess
    import tempfile
    from test_particle_amd import h5
    rng = np.random.default_rng(seed)
    planes = rng.standard_normal((6, n, n, n), dtype=np.float32)
    gb = planes.nbytes / 1e9
    lo, hi = (0.0, 0.0, 0.0), (float(n),) * 3
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        fname = os.path.join(tmp, "fields_0.h5")
        h5.write_pic_planes(fname, "Timestep_0", planes)
        del planes
        for rep in range(2):
            t0 = time.perf_counter()
            h5.load_pic(tr, fname, "Timestep_0", (n, n, n), lo, hi, 0.3, 1.7)
            w = time.perf_counter() - t0
        name = "fieldload_384_h5_file"
        out[name] = {"value": gb / w, "unit": "GB/s of datasets", "ms": w * 1e3,
                     "nodes": n ** 3}
        log(f"[bench] {name}: {gb / w:.2f} GB/s ({w * 1e3:.1f} ms)")
        exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
        if with_cpu and os.path.exists(exe):
            with open(os.path.join(tmp, "spec.txt"), "w") as f:
                f.write(f"system 4\nnx {n}\nny {n}\nnz {n}\ntmax 1.0\ndt 0.01\n"
                        f"gda_b0 1.7\ngda_e0 0.3\n")
                for a in ("x", "y", "z"):
                    f.write(f"{a}min 0.0\n{a}max {float(n)!r}\n")
            p = subprocess.run([exe, "h5", "spec.txt", fname, "Timestep_0", "-"], cwd=tmp,
                               check=True, capture_output=True, text=True, timeout=600)
            sec = float(dict(line.split() for line in p.stdout.splitlines()
                             if line.startswith("h5_"))["h5_seconds"])
            out[name]["cpu_baseline"] = {
                "value": gb / sec, "unit": "GB/s of datasets", "cores": 1, "kind": "reference",
                "sample": f"the same file through the reference's read_pic_efields_h5 and "
                          f"read_pic_bfields_h5 (oracle/_ref/ref_harness mode h5, one host "
                          f"thread as the reference runs them) on {cpu_model()}: "
                          f"{sec * 1e3:.0f} ms"}
            log(f"[bench] {name} CPU (reference loaders) {gb / sec:.2f} GB/s "
                f"({sec * 1e3:.0f} ms)")
    return out


def forms_openbc(args):
    """Open boundaries (the reference's bc_flag = 1) on the PIC-grid
    workload: the 256^3 grid, 1M particles at |v| = 0.7c, Boris, 5000
    steps, particles that leave the box stop and are counted in the escape
    spectrum.  value = particle-steps actually taken (the sum of the step
    counts) per second of wall time, beside the fraction remaining at the
    end and the rate of the same run with periodic boundaries."""
    from test_particle_amd import abi
    from test_particle_amd.fields import gyroperiod, pic_turbulence_grid
    from test_particle_amd.gpu import Tracker
    from test_particle_amd.particles import init_particles
    e, b, lo, hi = pic_turbulence_grid(args.grid, seed=args.seed)
    n, K = args.particles, 5000
    ptl = init_particles(n, lo, hi, 0.7, seed=args.seed)
    dt = 0.01 * gyroperiod(0.7, 1.0)
    out = {}
    with Tracker(0) as tr:
        tr.set_field_grid(e, b, lo, hi, abi.TP_GATHER_REF_COMPAT)
        for prec, name in ((abi.TP_FP64, "pic_grid_open_1M_fp64"),
                           (abi.TP_FP32_MIXED, "pic_grid_open_1M_fp32")):
            def prm(bc, steps=K):
                return abi.make_params(dt=dt, estep=steps, lo=lo, hi=hi,
                                       pusher=abi.TP_PUSH_BORIS, bc_flag=bc, precision=prec,
                                       nbins=102, nt_out=50, emin=1e-4, emax=1.0)
            tr.load(ptl)
            tr.run(prm(abi.TP_BC_OPEN, 100))
            tr.load(ptl)
            t0 = time.perf_counter()
            tr.run(prm(abi.TP_BC_OPEN))
            w = time.perf_counter() - t0
            _, nsteps = tr.store()
            d = tr.diagnostics(prm(abi.TP_BC_OPEN))
            taken = int(nsteps.astype(np.int64).sum())
            remain = float(d["remain"][-1]) / n
            escaped = int(d["espect_escape"].sum())
            tr.load(ptl)
            t1 = time.perf_counter()
            tr.run(prm(abi.TP_BC_PERIODIC))
            wp = time.perf_counter() - t1
            out[name] = {"value": taken / w, "unit": "particle-steps/s (steps taken)",
                         "dtype": "f64" if prec == abi.TP_FP64 else
                         "f32 push and gather, f64 positions",
                         "particles": n, "steps": K, "remaining_fraction": remain,
                         "escape_spectrum_counts": escaped,
                         "periodic_same_run_value": n * K / wp}
            log(f"[bench] {name}: {taken / w:.4e} particle-steps/s taken, remaining "
                f"{remain:.3f}, periodic {n * K / wp:.4e}")
    del e, b
    return out


def forms_grid2(args):
    """PIC fields with two time slices (the reference's multi_tframe = 1):
    the PIC-grid workload with a second, independent 256^3 slice, the field
    blended in time between them (slice spacing 20 time units, so the 1000
    timed steps cross several blend intervals); FP64 and FP32 Boris, 1M
    particles.  Algorithmic bytes: 384 per particle-step (two slices)."""
    from test_particle_amd import abi
    from test_particle_amd.fields import gyroperiod, pic_turbulence_grid
    from test_particle_amd.gpu import Tracker
    from test_particle_amd.particles import init_particles
    e1, b1, lo, hi = pic_turbulence_grid(args.grid, seed=args.seed)
    e0, b0, _, _ = pic_turbulence_grid(args.grid, seed=args.seed + 1)
    n, K = args.particles, 1000
    ptl = init_particles(n, lo, hi, 0.7, seed=args.seed)
    dt = 0.01 * gyroperiod(0.7, 1.0)
    out = {}
    with Tracker(0) as tr:
        tr.set_field_grid2(e1, b1, e0, b0, lo, hi, 20.0)
        for prec, name in ((abi.TP_FP64, "pic_grid2_1M_fp64"), (abi.TP_FP32_MIXED,
                                                                "pic_grid2_1M_fp32")):
            prm = abi.make_params(dt=dt, estep=K, lo=lo, hi=hi, pusher=abi.TP_PUSH_BORIS,
                                  precision=prec, nbins=102, nt_out=10, emin=1e-4, emax=1.0)
            tr.load(ptl)
            tr.run(abi.make_params(dt=dt, estep=100, lo=lo, hi=hi, pusher=abi.TP_PUSH_BORIS,
                                   precision=prec, nbins=102, nt_out=1, emin=1e-4,
                                   emax=1.0))
            tr.load(ptl)
            t0 = time.perf_counter()
            tr.run(prm)
            w = time.perf_counter() - t0
            kms, launches = tr.last_kernel_time()
            out[name] = {"value": n * K / w, "dtype": "f64" if prec == abi.TP_FP64 else
                         "f32 push and gather, f64 positions", "particles": n, "steps": K,
                         "algorithmic_GBps": 384.0 * n * K / (kms * 1e-3) / 1e9,
                         "kernel_ms_per_launch": kms / max(launches, 1)}
            log(f"[bench] {name}: {n * K / w:.4e} particle-steps/s, "
                f"{out[name]['algorithmic_GBps']:.0f} GB/s algorithmic")
        if not args.no_cpu:
            # the oracle (FP64, OpenMP) on the first 65,536 particles for 200
            # steps; the GPU runs the same sample and must match bit for bit
            from oracle import tpref
            cores = host_cores()
            f = tpref.Field.grid2(e1, b1, e0, b0, lo, hi, 20.0)
            ns, ks = 65536, 200
            prm = abi.make_params(dt=dt, estep=ks, lo=lo, hi=hi, pusher=abi.TP_PUSH_BORIS,
                                  nbins=102, nt_out=2, emin=1e-4, emax=1.0)
            ref = tpref.track_fixed(f, prm, ptl[:ns], nthreads=cores)
            g = tr.track_fixed(prm, ptl[:ns])
            same = bool(np.array_equal(g["ptl"], ref["ptl"]))
            out["pic_grid2_1M_fp64"]["cpu_baseline"] = {
                "value": ns * ks / ref["seconds"], "unit": "particle-steps/s", "cores": cores,
                "kind": "port",
                "sample": f"first {ns} particles, {ks} steps, same two slices, FP64 oracle "
                          f"with OpenMP ({cpu_model()})",
                "gpu_sample_bitwise_equal": same}
            log(f"[bench] pic_grid2 CPU {ns * ks / ref['seconds']:.3e} particle-steps/s on "
                f"{cores} threads; GPU sample bitwise={same}")
    del e1, b1, e0, b0
    return out


def forms_poincare(args, with_cpu=True):
    """Poincare map of the PIC-grid workload's magnetic field (the
    reference's run_type 2): a plane z = 128.3 of side x side start points,
    RK4 on B/|B| with step 0.05 cells and the reference's float trilinear
    gather, the reference's crossing test, `ncross` crossings per line,
    length cap 2e4.  A step is one RK4 step of one field line (four field
    evaluations); the value is total steps / wall time of the library call.
    The CPU baseline is the reference's own tracing on the first rows of the
    same plane (the oracle where the reference harness is not built), and
    the GPU's crossing points on those lines are compared bit for bit."""
    from test_particle_amd import abi
    from test_particle_amd.fields import pic_turbulence_grid
    from test_particle_amd.gpu import Tracker, pmap_starts
    _, b, lo, hi = pic_turbulence_grid(args.grid, seed=args.seed)
    side, nc = args.pmap_side, args.pmap_ncross
    st = pmap_starts(2, (0.0, 0.0, 128.3), min_h=lo[0] + 0.25, max_h=hi[0] - 0.25,
                     min_v=lo[1] + 0.25, max_v=hi[1] - 0.25, nh=side, nv=side)
    prm = abi.make_pmap_params(face=2, ncross=nc, step=0.05, length_max=2e4, lo=lo, hi=hi,
                               mode=abi.TP_PMAP_REF_COMPAT)
    out = {}
    e0 = np.zeros_like(b)
    with Tracker(0) as tr:
        tr.set_field_grid(e0, b, lo, hi, abi.TP_GATHER_REF_COMPAT)
        del e0
        warm = abi.make_pmap_params(face=2, ncross=1, step=0.05, length_max=50.0, lo=lo, hi=hi)
        tr.poincare_map(warm, st[:65536])
        t0 = time.perf_counter()
        r = tr.poincare_map(prm, st)
        w = time.perf_counter() - t0
        kms, launches = tr.last_kernel_time()
        steps = int(r["nsteps"].sum())
        name = f"poincare_{side}x{side}_lines_fp64"
        out[name] = {"value": steps / w, "unit": "field-line RK4 steps/s", "dtype": "f64",
                     "lines": int(st.shape[0]), "ncross": nc, "rk4_steps": steps,
                     "kernel_ms": kms, "launches": launches,
                     "kernel_rate": steps / (kms * 1e-3),
                     "field_evals_per_s": 4.0 * steps / w,
                     "lines_reaching_ncross": int((r["ncrossed"] == nc).sum())}
        log(f"[bench] {name}: {steps / w:.4e} RK4 steps/s ({steps} steps, {w:.2f}s, "
            f"{launches} launches)")
        # FP32 form: the same lines with the field, unit vectors and RK4 sum
        # in FP32 (positions, length and crossing test in FP64)
        prm32 = abi.make_pmap_params(face=2, ncross=nc, step=0.05, length_max=2e4, lo=lo,
                                     hi=hi, mode=abi.TP_PMAP_REF_COMPAT,
                                     precision=abi.TP_FP32_MIXED)
        t0 = time.perf_counter()
        r32 = tr.poincare_map(prm32, st)
        w32 = time.perf_counter() - t0
        k32, l32 = tr.last_kernel_time()
        s32 = int(r32["nsteps"].sum())
        both = (r["ncrossed"] >= 2) & (r32["ncrossed"] >= 2)
        dist = np.sqrt(sum((r32[k][both, 2] - r[k][both, 2]) ** 2
                           for k in ("xpos", "ypos", "zpos")))
        name32 = f"poincare_{side}x{side}_lines_fp32"
        out[name32] = {"value": s32 / w32, "unit": "field-line RK4 steps/s",
                       "dtype": "f32 field, unit vector and RK4 sum; f64 positions",
                       "rk4_steps": s32, "kernel_ms": k32, "launches": l32,
                       "vs_fp64_form": (s32 / w32) / (steps / w),
                       "same_crossing_count": float(np.mean(r32["ncrossed"] == r["ncrossed"])),
                       "second_record_dist_median": float(np.median(dist)),
                       "second_record_dist_p95": float(np.quantile(dist, 0.95))}
        log(f"[bench] {name32}: {s32 / w32:.4e} RK4 steps/s ({w32:.2f}s); same crossing "
            f"count {out[name32]['same_crossing_count']:.4f}, 2nd record distance median "
            f"{np.median(dist):.2e} p95 {np.quantile(dist, 0.95):.2e} cells")
        if with_cpu:
            base = pmap_reference_baseline(b, lo, hi, prm, side, r)
            if base is None:
                base = pmap_port_baseline(b, lo, hi, prm, st, r)
            out[name]["cpu_baseline"] = base
            log(f"[bench] {name} CPU ({base['kind']}) {base['value']:.3e} RK4 steps/s on "
                f"{base['cores']} threads; GPU sample bitwise={base['gpu_sample_bitwise_equal']}")
    return out


def abc_reference_baseline(ptl, dt, steps, d_gpu):
    """The reference's own tracking loop on the whole 16,384-particle x
    10,000-step Wirz workload: particle_tracking_hybrid compiled from its
    sources with A = B = C = 1 (oracle/_ref/ref_harness_abc1, mode hybrid,
    OpenMP over the box's threads), timed around the call (which includes
    the reference's own writes of its two small result files).  Its
    trajectories differ from the GPU's only through libm's sin/cos against
    the project's (ulps, amplified by the chaotic ABC flow), so D_xx is
    compared statistically.  None when the harness is absent."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness_abc1")
    if not os.path.exists(exe):
        return None
    cores = host_cores()
    two_pi = 2.0 * np.pi
    with tempfile.TemporaryDirectory() as tmp:
        os.makedirs(os.path.join(tmp, "data"))
        with open(os.path.join(tmp, "spec.txt"), "w") as f:
            f.write(f"system 2\nnx 5\nny 5\nnz 5\ntmax {float(steps * dt + 1e-9)!r}\n"
                    f"dt {float(dt)!r}\n"
                    f"bc 0\nnbins 102\nnt_out 100\nemin 0.001\nemax 10.0\n")
            for k in ("x", "y", "z"):
                f.write(f"{k}min 0.0\n{k}max {two_pi!r}\n")
        np.ascontiguousarray(ptl, np.float64).tofile(os.path.join(tmp, "p.bin"))
        p = subprocess.run([exe, "hybrid", "spec.txt", "p.bin"], cwd=tmp, check=True,
                           capture_output=True, text=True, timeout=600,
                           env=dict(os.environ, OMP_NUM_THREADS=str(cores)))
        kv = dict(line.split() for line in p.stdout.splitlines() if line.startswith("hybrid_"))
        sec = float(kv["hybrid_seconds"])
        diff = np.loadtxt(os.path.join(tmp, "data", "diff_coeffs.dat"), skiprows=1)
    dxx_ref = diff[-1, 1] / diff[-1, 8]
    dxx_gpu = d_gpu["dcoeff"][1, -1] / d_gpu["remain"][-1]
    return {"value": ptl.shape[0] * steps / sec, "unit": "particle-steps/s", "cores": cores,
            "kind": "reference",
            "sample": f"the whole workload ({ptl.shape[0]} particles x {steps} Wirz steps) "
                      f"through the reference's particle_tracking_hybrid "
                      f"(oracle/_ref/ref_harness_abc1, OpenMP) on {cpu_model()}",
            "dxx_last_rel_diff_vs_gpu": float(abs(dxx_gpu / dxx_ref - 1.0))}


def pmap_reference_baseline(b, lo, hi, prm, side, r):
    """The reference's own Poincare tracing -- its config_poincare_map and
    next_location compiled from its sources into oracle/_ref/ref_harness
    (mode pmap: the lines spread over OpenMP threads as in
    src/poincare_map.c:122) -- on the first rows of the same plane of start
    points, sized from a one-row pilot to about 12 s; its crossing points are
    compared with the GPU's bit for bit.  None when the harness is absent."""
    import subprocess
    import tempfile
    from test_particle_amd.poincare import write_config
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(exe):
        return None
    cores = host_cores()
    dv = (hi[1] - 0.25 - (lo[1] + 0.25)) / (side - 1)
    with tempfile.TemporaryDirectory() as tmp:
        grid = os.path.join(tmp, "grid.bin")
        with open(grid, "wb") as f:
            np.zeros(b.size, np.float32).tofile(f)          # E is not used
            np.ascontiguousarray(b, np.float32).tofile(f)
        nz, ny, nx = b.shape[:3]
        with open(os.path.join(tmp, "spec.txt"), "w") as f:
            f.write(f"system 4\nnx {nx}\nny {ny}\nnz {nz}\ntmax 1.0\ndt 0.01\ngrid {grid}\n")
            for k, a in zip(("x", "y", "z"), range(3)):
                f.write(f"{k}min {float(lo[a])!r}\n{k}max {float(hi[a])!r}\n")

        def run(rows):
            cfg = dict(face=2, step=prm.step, length_max=prm.length_max, ncross=prm.ncross,
                       single=0, x0=0.0, y0=0.0, z0=128.3, min_h=lo[0] + 0.25,
                       max_h=hi[0] - 0.25, min_v=lo[1] + 0.25,
                       max_v=lo[1] + 0.25 + dv * (rows - 1), nh=side, nv=rows)
            write_config(os.path.join(tmp, "init.dat"), cfg)
            p = subprocess.run([exe, "pmap", "spec.txt", "init.dat", "pm.bin"], cwd=tmp,
                               check=True, capture_output=True, text=True, timeout=300,
                               env=dict(os.environ, OMP_NUM_THREADS=str(cores)))
            kv = dict(line.split() for line in p.stdout.splitlines()
                      if line.startswith("pmap_"))
            out = np.fromfile(os.path.join(tmp, "pm.bin")).reshape(3, side * rows, -1)
            return float(kv["pmap_seconds"]), int(kv["pmap_steps"]), out

        sec, steps, _ = run(1)
        rows = int(np.clip(12.0 / max(sec, 1e-3), 1, side))
        sec, steps, out = run(rows)
        m = side * rows
        same = all(np.array_equal(out[k], r[key][:m])
                   for k, key in enumerate(("xpos", "ypos", "zpos")))
    return {"value": steps / sec, "unit": "field-line RK4 steps/s", "cores": cores,
            "kind": "reference",
            "sample": f"first {m} lines ({rows} rows) of the same plane: the reference's "
                      f"next_location and config parser (oracle/_ref/ref_harness pmap, "
                      f"OpenMP over lines) on {cpu_model()}",
            "gpu_sample_bitwise_equal": bool(same)}


def pmap_port_baseline(b, lo, hi, prm, st, r):
    """The CPU oracle (OpenMP) on the first lines of the plane, about 12 s."""
    from oracle import tpref
    from test_particle_amd import abi
    cores = host_cores()
    f = tpref.Field.grid(np.zeros_like(b), b, lo, hi, abi.TP_GATHER_REF_COMPAT)
    pilot = tpref.poincare_map(f, prm, st[:4 * cores], nthreads=cores)
    rate = pilot["nsteps"].sum() / max(pilot["seconds"], 1e-6)
    m = int(np.clip(12.0 * rate / pilot["nsteps"].mean(), 4 * cores, 65536))
    ref = tpref.poincare_map(f, prm, st[:m], nthreads=cores)
    same = all(np.array_equal(r[k][:m], ref[k]) for k in ("xpos", "ypos", "zpos"))
    return {"value": ref["nsteps"].sum() / ref["seconds"], "unit": "field-line RK4 steps/s",
            "cores": cores, "kind": "port",
            "sample": f"first {m} lines of the same plane, FP64 oracle with OpenMP "
                      f"({cpu_model()})",
            "gpu_sample_bitwise_equal": bool(same)}


def cpu_baseline(e, b, lo, hi, ptl, params, tr):
    """The CPU oracle (FP64, OpenMP) on this host's cores, on a bounded sample
    of the same workload: the first particles of the same set for a number of
    steps chosen from a short pilot so the timed sample is about 10-20 s.  The
    GPU runs the same sample, and the two are compared."""
    from oracle import tpref
    from test_particle_amd import abi
    cores = host_cores()
    tpref.set_math(False)
    f = tpref.Field.grid(e, b, lo, hi, abi.TP_GATHER_REF_COMPAT)
    n_s = min(ptl.shape[0], 65536)
    sample = ptl[:n_s]
    pilot = tpref.track_fixed(f, params(20), sample[:8192], nthreads=cores)
    rate = 8192 * 20 / max(pilot["seconds"], 1e-6)
    steps = int(np.clip(12.0 * rate / n_s, 100, 2000)) // 100 * 100
    log(f"[bench] CPU baseline: {n_s} particles x {steps} steps on {cores} threads "
        f"(pilot {rate:.3e}/s)")
    ref = tpref.track_fixed(f, params(steps), sample, nthreads=cores)
    cpu_rate = n_s * steps / ref["seconds"]
    g = tr.track_fixed(params(steps), sample)
    same = bool(np.array_equal(g["ptl"], ref["ptl"]))
    dx_ref = ref["dcoeff"][1, 1:]
    dx_gpu = g["dcoeff"][1, 1:]
    derr = float(np.max(np.abs(dx_gpu - dx_ref) / np.abs(dx_ref)))
    log(f"[bench] CPU {cpu_rate:.3e} particle-steps/s; GPU sample bitwise={same}, "
        f"D_xx rel err {derr:.1e}")
    return {"value": cpu_rate, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": f"first {n_s} of the same particles, {steps} steps, same grid and "
                      f"params, FP64 oracle with OpenMP ({cpu_model()})",
            "gpu_sample_bitwise_equal": same, "gpu_sample_dxx_rel_err": derr}


if __name__ == "__main__":
    main()
