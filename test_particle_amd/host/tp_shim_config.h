/*
 * tp_shim_config.h -- the optional "GPU tracking" section of init.dat read by
 * the drop-in shim (tp_hybrid_shim.c).
 *
 * The reference program reads its settings from the fixed file init.dat
 * (src/main.c:30), one section at a time: a line holding the section's title
 * is found with strstr, then fixed "Label: value" lines follow
 * (src/domain.c:86-157, src/particle_info.c:114-179).  The GPU section uses
 * the same layout; every line in it is optional and the section itself may
 * be absent (the reference's parser skips lines it does not look for, so the
 * section can sit anywhere in the file):
 *
 *   ------------------------------- GPU tracking -------------------------------
 *   Pusher (wirz or boris): boris
 *   Precision (fp64 or fp32): fp32
 *   Check for non-finite states (0 or 1): 1
 *   Fix stale diffusion terms of escaped particles (0 or 1): 0
 *   Adaptive relative tolerance: 1e-14
 *   Adaptive absolute tolerance: 1e-5
 *   Adaptive first step: 1e-5
 *   Adaptive maximum sequence index: 8
 *   ------------------------------------------------------------------------------
 *
 * The section ends at the next line that starts with "---" or at the end of
 * the file.  Defaults are the reference's own choices: the Wirz pusher in
 * double precision (src/tracking.c:176-180), the Bulirsch-Stoer constants
 * KMAXX 8, atol 1e-5, rtol 1e-14 (include/StepperBS.h:18-22) and first step
 * 1e-5 (src/tracking.c:566).
 *
 * Plain C over stdio; no MPI, HIP or reference headers, so it builds and is
 * tested on its own.
 */
#ifndef TP_SHIM_CONFIG_H
#define TP_SHIM_CONFIG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int found;          /* 1 if init.dat holds a GPU tracking section */
    int pusher;         /* TP_PUSH_WIRZ or TP_PUSH_BORIS */
    int precision;      /* TP_FP64 or TP_FP32_MIXED */
    int flags;          /* TP_FLAG_CHECK_FINITE | TP_FLAG_FIX_STALE_DCOEFF */
    double bs_rtol, bs_atol, bs_h0;
    int bs_kmax;
} tp_shim_config;

/* the reference's choices (see above) */
void tp_shim_config_defaults(tp_shim_config *c);

/* Reads the GPU tracking section of the file at path into c (defaults
 * first).  A missing file or section leaves the defaults and returns 0.  A
 * line of the section that is not one of the labels above, or a value out of
 * range, returns -1 with a message in err (at most nerr bytes, with the line
 * number).  Precision fp32 needs pusher boris (the Wirz pusher has no FP32
 * form). */
int tp_shim_config_read(const char *path, tp_shim_config *c, char *err, size_t nerr);

#ifdef __cplusplus
}
#endif

#endif
