"""Achievable peaks of the GPU next to MI355X's published ones.

    python -m test_particle_amd.peaks [--reps 5] [--gib 4]

Runs the microbenchmarks of tp_peak_measure (include/tp_peak.h): an HBM copy
and a read-only sweep (16 bytes per lane per access), FP64 and FP32 FMA
chains (16 per lane; FP32 unpacked, as this library is built with packed FP32
math disabled) and a bf16 MFMA loop (v_mfma_f32_32x32x16_bf16, 4 accumulators
per wave), and prints one JSON line with each rate, the fastest launch's
duration and the published peak it compares with.
"""
import argparse
import json

from . import abi

# MI355X published peaks: HBM3E 8 TB/s; dense FP64 / FP32 vector and bf16
# MFMA rates in TFLOP/s
PUBLISHED = {"hbm_copy": (8000.0, "GB/s"), "hbm_read": (8000.0, "GB/s"),
             "fp64_fma": (78.6, "TFLOP/s"), "fp32_fma": (157.3, "TFLOP/s"),
             "bf16_mfma": (2500.0, "TFLOP/s")}
KINDS = {"hbm_copy": abi.TP_PEAK_HBM_COPY, "hbm_read": abi.TP_PEAK_HBM_READ,
         "fp64_fma": abi.TP_PEAK_FP64_FMA, "fp32_fma": abi.TP_PEAK_FP32_FMA,
         "bf16_mfma": abi.TP_PEAK_BF16_MFMA}


def measure(reps=5, nbytes=0, device=0, kinds=None):
    from .gpu import measure_peak
    out = {}
    for name in kinds or KINDS:
        v, ms = measure_peak(KINDS[name], nbytes, reps, device)
        peak, unit = PUBLISHED[name]
        out[name] = {"value": v, "unit": unit, "ms": ms, "published": peak,
                     "frac_of_published": v / peak}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--gib", type=float, default=4.0, help="HBM sweep size per buffer")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args()
    r = measure(a.reps, int(a.gib * (1 << 30)), a.device)
    for name, x in r.items():
        print(f"{name}: {x['value']:.1f} {x['unit']} ({x['ms']:.3f} ms per launch), "
              f"{100 * x['frac_of_published']:.1f} % of the published {x['published']}",
              flush=True)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
