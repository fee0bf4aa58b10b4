"""Fixed workloads for profiling runs (rocprofv3) of the tracking kernels.

    python -m test_particle_amd.profile_run grid  [--steps 1000]   # PIC-grid tracking
    python -m test_particle_amd.profile_run abc   [--steps 1000]   # ABC field tracking
    python -m test_particle_amd.profile_run copy  [--mib 1024]     # plain copy kernel

``copy`` launches one device-to-device copy kernel (torch's) of a known size;
its FETCH_SIZE / WRITE_SIZE readings calibrate those counters against bytes.
"""
import argparse
import sys
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["grid", "abc", "copy"])
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--particles", type=int, default=1 << 20)
    ap.add_argument("--pusher", default="boris", choices=["boris", "wirz"])
    ap.add_argument("--fp32", action="store_true")
    ap.add_argument("--mib", type=int, default=1024)
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
    pusher = abi.TP_PUSH_BORIS if a.pusher == "boris" else abi.TP_PUSH_WIRZ
    prec = abi.TP_FP32_MIXED if a.fp32 else abi.TP_FP64
    if a.what == "grid":
        e, b, lo, hi = pic_turbulence_grid(256)
        tr.set_field_grid(e, b, lo, hi, abi.TP_GATHER_REF_COMPAT)
        dt = 0.01 * gyroperiod(0.7, 1.0)
    else:
        lo, hi = (0.0, 0.0, 0.0), (2 * np.pi,) * 3
        tr.set_field_abc(1.0, 1.0, 1.0)
        dt = 0.01 * gyroperiod(0.7, abc_b_ref())
    ptl = init_particles(a.particles, lo, hi, 0.7)
    tr.load(ptl)
    prm = abi.make_params(dt=dt, estep=a.steps, lo=lo, hi=hi, pusher=pusher, precision=prec,
                          nbins=102, nt_out=max(1, a.steps // 100), emin=1e-4, emax=1.0)
    t0 = time.perf_counter()
    tr.run(prm)
    wall = time.perf_counter() - t0
    ms, k = tr.last_kernel_time()
    print(f"{a.what}: {a.particles} x {a.steps} steps, {a.particles * a.steps / wall:.4e} "
          f"particle-steps/s, kernels {ms:.2f} ms / {k} launches", flush=True)
    tr.close()
    sys.stdout.flush()


if __name__ == "__main__":
    main()
