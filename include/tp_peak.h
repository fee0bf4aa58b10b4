/*
 * tp_peak.h -- achievable peaks of the GPU, measured by microbenchmarks.
 *
 * No counterpart in the reference.  The roofline fractions of the step
 * kernels are quoted against MI355X's published peaks; these kernels measure
 * what the same device reaches on plain streaming and arithmetic loops, so
 * that both figures sit side by side:
 *
 *   TP_PEAK_HBM_COPY   GB/s (bytes read + bytes written) of a 16-byte-per-lane
 *                      device-to-device copy of `bytes` bytes
 *   TP_PEAK_HBM_READ   GB/s of a read-only sweep over `bytes` bytes
 *   TP_PEAK_FP64_FMA   TFLOP/s of v_fma_f64 chains (2 FLOP per FMA)
 *   TP_PEAK_FP32_FMA   TFLOP/s of v_fma_f32 chains; this library is built with
 *                      packed FP32 math disabled, so this is the unpacked rate
 *   TP_PEAK_BF16_MFMA  TFLOP/s of v_mfma_f32_32x32x16_bf16 (dense, 32768 FLOP
 *                      per instruction and wave)
 *
 * tp_peak_measure runs the kernel `reps` times after one untimed launch (the
 * HBM kernels at each of 4, 16 and 64 blocks of 256 lanes per CU) and returns
 * the rate of the fastest launch and its duration (HIP events on a stream of
 * its own).  `bytes` is used by the HBM kernels (0: 4 GiB) and
 * ignored by the others.  Allocates and frees its own device memory.
 * Returns TP_OK, TP_E_ARG, TP_E_NOMEM, TP_E_HIP or TP_E_NODEV.
 */
#ifndef TP_PEAK_H
#define TP_PEAK_H

#include <stddef.h>
#include "tp_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define TP_PEAK_HBM_COPY   0
#define TP_PEAK_HBM_READ   1
#define TP_PEAK_FP64_FMA   2
#define TP_PEAK_FP32_FMA   3
#define TP_PEAK_BF16_MFMA  4

int tp_peak_measure(int device, int what, size_t bytes, int reps, double *value,
                    double *ms);

#ifdef __cplusplus
}
#endif

#endif /* TP_PEAK_H */
