/*
 * tpref.h -- the project's FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C restatement of the reference tracer's hot path, used by tests/
 * as the parity checker, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg.  The product (test_particle_amd/, libtpgpu.so) never
 * links, loads or calls it.  Each function names the reference lines it
 * follows; see the pins listed in oracle/tpref.c.
 */
#ifndef TPREF_H
#define TPREF_H

#include "../include/tp_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define TPREF_FIELD_TEST  0
#define TPREF_FIELD_WLCS  1
#define TPREF_FIELD_ABC   2
#define TPREF_FIELD_GRID  4
#define TPREF_FIELD_MODES 5
#define TPREF_FIELD_MHD   3   /* system 3: float v (egrid) and B (bgrid), E = B x v */
#define TPREF_FIELD_GRIDD 6   /* PIC grid of double E and B (egrid_d, bgrid_d) */

typedef struct tpref_field {
    int kind;
    /* ABC: A, B, C, lambda, B0 */
    double abc[5];
    /* optional velocity grid for E = B x v (ABC) */
    int vnx, vny, vnz;
    double vlo[3], vhi[3];
    const double *vgrid;
    /* PIC grid */
    int nx, ny, nz;
    double lo[3], hi[3];
    const float *egrid, *bgrid;
    int gather_mode;
    /* multi-mode */
    int nmodes;
    const double *modes;
    double b0z;
    /* WLCS */
    int nwlcs;
    const double *wlcs_rows;
    int ibessel;
    double omega, norm_i;
    int wlcs_mode;          /* TP_GATHER_REF_COMPAT: the reference's ellint */
    /* PIC grid, second time slice (multi_tframe = 1): egrid/bgrid hold the
     * next slice (the reference's efa/maga), egrid_b/bgrid_b the previous
     * one (efb/magb), grid_dt the time between slices */
    const float *egrid_b, *bgrid_b;
    double grid_dt;
    /* double E and B nodes (TPREF_FIELD_GRIDD), double3 AoS, x fastest */
    const double *egrid_d, *bgrid_d;
} tpref_field;

/* 1: take sin/cos from libm (as the reference does); 0: tp_sincos, the
 * routine the GPU kernels use. */
void tpref_set_math(int use_libm);
/* reproduce the reference's single-thread carry-over of stale D terms */
void tpref_set_carry_stale(int on);

void tpref_eval_field(const tpref_field *f, int n, const double *pts,
                      double *out);
void tpref_wirz_step(const tpref_field *f, tp_particle *p, double dt,
                     double qm);
void tpref_boris_step(const tpref_field *f, tp_particle *p, double dt,
                      double qm);
void tpref_particle_bc(const tp_track_params *prm, tp_particle *p,
                       int *iescape);
int  tpref_energy_bin(const tp_particle *p, int nbins, double logemin,
                      double logde);
void tpref_diff_coeff(const tpref_field *f, const tp_particle *p0,
                      const tp_particle *p, double dt, double out[8]);
/* Returns 0 or TP_E_ARG.  nthreads <= 0 uses the OpenMP default.
 * *seconds (if given) receives the wall time of the step loop. */
int  tpref_track_fixed(const tpref_field *f, const tp_track_params *prm,
                       int n, tp_particle *ptl, const tp_particle *ptl_init,
                       int *nsteps, double *espect, double *espect_escape,
                       double *dcoeff, int *nptl_remain, int nthreads,
                       double *seconds);
/* adaptive Bulirsch-Stoer of every particle to its end time; nsteps gets
 * accepted steps, totals = {accepted, rejected, rhs evaluations, failed} */
int  tpref_bs_track(const tpref_field *f, const tp_bs_params *prm, int n,
                    tp_particle *ptl, int *nsteps, long long *totals,
                    int nthreads, double *seconds);
/* the same with NR3 dense output, particle_bc after each accepted step and
 * the fixed-step loop's diagnostics at its output times (sstep must be 1;
 * the end time is t + estep*dt); with trajectory diagnostics armed
 * (tpref_set_trajectory) the particle at the output steps j = k
 * nsteps_output, from the dense output inside a step (folded into the box
 * for periodic boundaries) or the final state at the end time */
int  tpref_bs_track_diag(const tpref_field *f, const tp_bs_params *bs,
                         const tp_track_params *prm, int n, tp_particle *ptl,
                         const tp_particle *ptl_init, int *nsteps, double *espect,
                         double *espect_escape, double *dcoeff, int *nptl_remain,
                         long long *totals, int nthreads, double *seconds);
/* Poincare map of the field at t = 0 (src/poincare_map.c:39-258); arguments
 * as for tp_gpu_poincare_map.  Returns 0 or TP_E_ARG. */
int  tpref_poincare_map(const tpref_field *f, const tp_pmap_params *prm, int n,
                        const double *starts, double *xpos, double *ypos, double *zpos,
                        int *ncrossed, long long *nsteps, int nthreads, double *seconds);
/* exported for the tests of the shared routines */
void tpref_sincos(double x, double *s, double *c);
void tpref_ellint(double k, double *elle, double *ellk);
void tpref_ellint_ref(double k, double *elle, double *ellk);

#ifdef __cplusplus
}
#endif
#endif
