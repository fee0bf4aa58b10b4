/*
 * tp_ctx.h -- library-private context shared by the HIP translation units.
 */
#ifndef TP_CTX_H
#define TP_CTX_H

#include <hip/hip_runtime.h>
#include <stdio.h>
#include "../../include/tp_gpu.h"
#include "tp_device.h"

enum { PREC64 = 0, PREC32 = 1 };

/* ------------------------------------------------------------------ */
/* context                                                             */
/* ------------------------------------------------------------------ */
struct State {
    int cap = 0, n = 0;
    double *buf = nullptr;   /* all double arrays, 26 * cap */
    int *ibuf = nullptr;     /* esc, nsteps: 2 * cap */
    tp_particle *stage = nullptr;
    double *x, *y, *z, *ux, *uy, *uz, *t, *xs, *ys, *zs;
    double *x0, *y0, *z0, *ux0, *uy0, *uz0, *t0, *g0, *mu0, *th0;
    double *st;              /* 5 * cap stale D terms */
    int *esc, *nsteps;
    bool derived = false;    /* g0/mu0/th0 valid for current field */
};

struct tp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    FieldDev f;
    bool have_field = false;
    /* owned field buffers */
    float4 *grid = nullptr, *grid_b = nullptr;
    float4 *cells = nullptr, *cells_b = nullptr;   /* cell-major copies */
    int cells_mode = TP_CELLS_AUTO;
    double *vgrid = nullptr, *modes = nullptr;
    double *gridd = nullptr;                       /* double PIC nodes */
    float *modes32 = nullptr;
    WlcsSys *wl = nullptr;
    State s;
    /* diagnostics of the last run */
    int nb = 0, nt = 0, nd = 0;
    unsigned *espect = nullptr, *espect_esc = nullptr;
    double *dco = nullptr, *partials = nullptr;
    int *rdec = nullptr;
    size_t partials_cap = 0, diag_cap = 0, dco_cap = 0;
    /* open boundaries: indices of the particles still inside, and their count */
    int *live_idx = nullptr;
    size_t live_cap = 0;
    int last_n = 0;
    double last_ms = 0.0;
    int last_launches = 0;
    int num_cu = 0;
    hipEvent_t *evs = nullptr;   /* start/stop per step-kernel launch */
    int nev = 0;
    /* adaptive stepper: counters and the captured graph of its launch */
    unsigned long long *bs_ctr = nullptr;
    hipGraphExec_t bs_graph = nullptr;
    unsigned long long bs_graph_key = 0;
    int *bs_tune_block = nullptr;
    /* adaptive stepper with diagnostics: dense-output scratch per lane, D
     * terms per particle and row, counted-out histogram */
    double *bs_ds = nullptr, *bs_dbuf = nullptr;
    int *bs_outhist = nullptr;
    size_t bs_ds_cap = 0, bs_dbuf_cap = 0, bs_outhist_cap = 0;
    /* learned surrogate: network, optimiser state and training buffers */
    struct MlpState *mlp = nullptr;
    /* trajectory diagnostics armed for the next run (traj_nout > 0) */
    int traj_nout = 0;
    tp_particle *traj = nullptr;
    int *traj_accum = nullptr, *traj_pts = nullptr;
    size_t traj_cap = 0;
    int traj_ncap = 0;
    /* grid input: two pinned host slabs and their device copies, kept for
     * the next load (freed by tp_gpu_finalize) */
    float *stage_pin[2] = {nullptr, nullptr};
    float *stage_dev[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    size_t stage_floats = 0;
    /* device address of the static scratch array on this context's GPU
     * (tp_gpu_run_fixed_noalloc); null if the lookup failed */
    double *scratch = nullptr;
};

void tp_bs_release(tp_ctx *c);
void tp_mlp_release(tp_ctx *c);

#define HIPCHK(x)                                                        \
    do {                                                                 \
        hipError_t e_ = (x);                                             \
        if (e_ != hipSuccess) {                                          \
            fprintf(stderr, "tp_gpu: %s failed: %s (%s:%d)\n", #x,       \
                    hipGetErrorString(e_), __FILE__, __LINE__);          \
            return TP_E_HIP;                                             \
        }                                                                \
    } while (0)

/* ------------------------------------------------------------------ */
/* device helpers                                                      */
/* ------------------------------------------------------------------ */

static inline int field_kind(const tp_ctx *c) { return c->f.kind; }

#endif /* TP_CTX_H */
