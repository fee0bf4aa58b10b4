/*
 * tp_gpu.h -- C ABI of the MI355X test-particle tracking library
 * (libtpgpu.so, HIP kernels for gfx950).
 *
 * Drop-in point: the reference tracer's tracking loop
 *   void particle_tracking_hybrid(int mpi_rank, int nptl, double dt,
 *        int nbins, int nt_out, int bc_flag, int nsteps_output, double pmass,
 *        int *ntraj_accum, int tracking_method, int traj_diagnose,
 *        particles *ptl_init, int *nsteps_ptl_tracking, particles *ptl,
 *        particles *ptl_time);
 * (src/tracking.c:659-699, declared include/tracking.h:78-82).  The
 * reference passes the field, domain, charge/mass and spectrum settings
 * through file-scope globals; here they are explicit: a context holds the
 * field source (tp_gpu_set_field_*), and every call takes a
 * tp_track_params.  No function of this library calls exit(), reads or
 * writes files, or uses MPI; all return TP_OK (0) or a negative TP_E_*
 * code, and tp_gpu_strerror() names it.
 *
 * Threading: one context per host thread; calls on one context are
 * serialised by the caller.  Host buffers are read/written synchronously
 * (every call has finished with them when it returns).
 */
#ifndef TP_GPU_H
#define TP_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One particle; byte-for-byte the reference's struct particles
 * (include/particle_info.h:6-12): 11 doubles, 88 bytes. */
typedef struct tp_particle {
    double x, y, z, qm;
    double ux, uy, uz, t;     /* u = gamma * v / c */
    double xshift, yshift, zshift;
} tp_particle;

/* error codes */
#define TP_OK            0
#define TP_E_ARG        -1   /* bad argument or inconsistent parameters */
#define TP_E_NOMEM      -2   /* device or host allocation failed */
#define TP_E_HIP        -3   /* a HIP runtime call failed */
#define TP_E_NONFINITE  -4   /* NaN/Inf in particle state after tracking */
#define TP_E_STATE      -5   /* call out of order (no field, no particles) */
#define TP_E_NODEV      -6   /* no usable GPU */

/* pushers (the reference's tracking_wirz, src/tracking.c:749-892, and a
 * relativistic Boris push, which the reference does not have) */
#define TP_PUSH_WIRZ   0
#define TP_PUSH_BORIS  1

/* arithmetic of the push and field evaluation */
#define TP_FP64        0   /* everything in FP64 */
#define TP_FP32_MIXED  1   /* FP32 push/field, FP64 positions and diagnostics */

/* particle boundary conditions, src/tracking.c:1172-1221 */
#define TP_BC_PERIODIC 0
#define TP_BC_OPEN     1
#define TP_BC_NONE     2

/* trilinear gather conventions for gridded fields */
#define TP_GATHER_REF_COMPAT 0  /* src/interpolation.c:141-181 weights (x and
                                   z fractions exchanged), float rounding after
                                   each corner (src/magnetic_field.c:603-608),
                                   wrap cell of src/interpolation.c:61-90 */
#define TP_GATHER_TEXTBOOK   1  /* standard trilinear weights, FP64
                                   accumulation, periodic wrap */

/* tp_track_params.flags */
#define TP_FLAG_FIX_STALE_DCOEFF 1  /* zero the per-particle D terms each step
                                       (the reference keeps stale dxx/dyy/dzz/
                                       daa/dee, src/tracking.c:334-336) */
#define TP_FLAG_CHECK_FINITE     2  /* after the run, scan the particles' positions
                                       and momenta on the GPU and return
                                       TP_E_NONFINITE if any is NaN or infinite
                                       (the reference lets them propagate: B = 0
                                       makes iBtot infinite, src/tracking.c:782);
                                       the outputs are still written */

typedef struct tp_track_params {
    int pusher;            /* TP_PUSH_* */
    int precision;         /* TP_FP64 / TP_FP32_MIXED */
    int bc_flag;           /* TP_BC_* (src/tracking.c:61-65) */
    int nbins;             /* energy bins (include/Global.h:22) */
    int nt_out;            /* spectrum frames */
    int sstep, estep;      /* steps sstep..estep inclusive (src/tracking.c:333);
                              the reference uses sstep = 1 and
                              estep = (int)(tmax/dt) (src/tracking.c:103) */
    int einterval_t;       /* spectrum cadence; reference: estep/nt_out
                              (src/tracking.c:104) */
    int dcoeff_interval;   /* D cadence; reference: 100 (src/tracking.c:34) */
    int flags;             /* TP_FLAG_* */
    double dt;             /* time step */
    double qm;             /* charge/mass: the reference's global charge_mass
                              (src/tracking.c:787), not ptl->qm */
    double xmin, xmax, ymin, ymax, zmin, zmax;   /* simul_domain */
    double logemin, logde; /* src/diagnostics.c:629-631 */
} tp_track_params;

typedef struct tp_ctx tp_ctx;

/* --- context ------------------------------------------------------------ */
int  tp_gpu_init(int device, tp_ctx **ctx);
void tp_gpu_finalize(tp_ctx *ctx);
const char *tp_gpu_strerror(int code);
/* device name and compute-unit count of the context's GPU */
int  tp_gpu_device_info(tp_ctx *ctx, char *name, int name_len, int *num_cu);

/* --- field sources (replace get_emf's switch, src/emfields.c:32-60) ------ */
/* system 0: B = z-hat, E = 0 (getemf_test, src/emfields.c:210-218) */
int tp_gpu_set_field_test(tp_ctx *ctx);
/* system 2: ABC force-free B = B0 (A sin lz + C cos ly, B sin lx + A cos lz,
 * C sin ly + B cos lx) (getb_ff, src/force_free.c:138-150); E = B x v from
 * the velocity grid set by tp_gpu_set_velocity_grid, or 0 without one
 * (gete_ff, src/force_free.c:78-86). */
int tp_gpu_set_field_abc(tp_ctx *ctx, double A, double B, double C,
                         double lambda, double B0);
/* nodes (nx, ny, nz) over [lo, hi] endpoint-inclusive (dx = (hi-lo)/(n-1),
 * src/domain.c:548-550), x fastest, AoS double3 velocity
 * (get_double_velocity_at_point, src/velocity_field.c:555-584).
 * v == NULL removes it; nx*ny*nz must be below 2^32 - 1 (TP_E_ARG). */
int tp_gpu_set_velocity_grid(tp_ctx *ctx, int nx, int ny, int nz,
                             const double lo[3], const double hi[3],
                             const double *v);
/* system 4: PIC grids, AoS float3 E and B on the same nodes, x fastest
 * (getemf_pic_test_particle, src/emfields.c:153-166); nx*ny*nz below
 * 2^32 - 1 (TP_E_ARG), as for every grid loader below. */
int tp_gpu_set_field_grid(tp_ctx *ctx, int nx, int ny, int nz,
                          const double lo[3], const double hi[3],
                          const float *efield, const float *bfield,
                          int gather_mode);
/* Two time slices of the PIC fields (the reference's multi_tframe = 1,
 * src/interpolation.c:97-100, src/electric_field.c:407-437,
 * src/magnetic_field.c:593-622): at time t the field is
 *   dt * F_next + (1 - dt) * F_prev,  dt = t/grid_dt - floor(t/grid_dt),
 * each slice gathered with the same trilinear weights and summed in double
 * (the reference gather then rounds the blend to float).  F_next and F_prev
 * are what the reference loads as efa/maga (slice ct + 1) and efb/magb
 * (slice ct - 1) (src/magnetic_field.c:440-530); layouts as for
 * tp_gpu_set_field_grid. */
int tp_gpu_set_field_grid2(tp_ctx *ctx, int nx, int ny, int nz, const double lo[3],
                           const double hi[3], const float *efield_next,
                           const float *bfield_next, const float *efield_prev,
                           const float *bfield_prev, double grid_dt, int gather_mode);

/* system 3: MHD fields, AoS float3 velocity and B on the same nodes, x
 * fastest (getemf_mhd_test_particle, src/emfields.c:107-120): v and B are
 * gathered as the reference's float gathers do (get_float_velocity_at_point,
 * src/velocity_field.c:593-622, get_float_bfield_at_point,
 * src/magnetic_field.c:593-622, reference weights and per-corner float
 * rounding) and E = B x v is formed in double.  One time slice (the
 * reference's fixed-step loop does not track system 3 with two,
 * src/tracking.c:151-157); the cell-major copy applies as for PIC grids. */
int tp_gpu_set_field_mhd(tp_ctx *ctx, int nx, int ny, int nz, const double lo[3],
                         const double hi[3], const float *vfield, const float *bfield);
/* PIC grids of doubles, AoS double3 E and B, x fastest: the reference's
 * double gathers get_double_efield_at_point and get_double_bfield_at_point
 * (src/electric_field.c:370-406, src/magnetic_field.c:555-591; its weights,
 * double accumulators), one time slice.  The reference defines them but
 * never reaches them: its setup loads systems 3 and 4 as float grids
 * (src/domain.c:357-400) and get_emf calls the float gathers. */
int tp_gpu_set_field_grid_double(tp_ctx *ctx, int nx, int ny, int nz, const double lo[3],
                                 const double hi[3], const double *efield,
                                 const double *bfield);

/* Cell-major copy of a PIC grid.  Besides the nodes, the library can keep
 * every cell's eight corners (the reference gather's corners, six floats
 * each: 192 bytes per cell, 6x the nodes) so that a particle entering a cell
 * reads one contiguous record; the gathered values and all results are the
 * same either way.  TP_CELLS_AUTO (the default) builds the copy when it
 * takes at most half of the free device memory; TP_CELLS_OFF never;
 * TP_CELLS_ON always (TP_E_NOMEM if it does not fit).  The mode applies to
 * the grids set after the call. */
#define TP_CELLS_AUTO 0
#define TP_CELLS_OFF  1
#define TP_CELLS_ON   2
int tp_gpu_set_grid_cells(tp_ctx *ctx, int mode);

/* --- PIC field input pipeline ------------------------------------------
 * The reference's loaders read one component at a time as a plane of floats
 * and scatter it, scaled, into the float3 node arrays on the host
 * (read_efields_float_binary, src/electric_field.c:257-300, for .gda files,
 * x fastest; read_pic_efields_float_h5, src/electric_field.c:447-507, and
 * read_pic_bfields_float_h5, src/magnetic_field.c:632-692, for HDF5, x
 * slowest and transposed).  Here the raw planes go to the GPU in slabs along
 * their slowest axis through two pinned staging buffers (the host fills one
 * while the other is copied and laid out), and a GPU kernel scales, transposes
 * and interleaves them into the nodes: node values are (float)(v * e0) and
 * (float)(v * b0) in double, as in the reference, so the field is the same as
 * tp_gpu_set_field_grid / _grid2 of the reference's arrays. */
#define TP_PLANES_X_FASTEST 0   /* plane index (k*ny + j)*nx + i: .gda files */
#define TP_PLANES_X_SLOWEST 1   /* plane index (i*ny + j)*nz + k: PIC HDF5 */

/* Source of the planes: writes component comp (0..5 = ex ey ez bx by bz) of
 * slice (0 = the next or only slice, 1 = the previous slice) for slowest-axis
 * indices [s0, s1) to dst, (s1 - s0) * (the product of the two other
 * dimensions) floats in the plane's order.  Returns 0, or nonzero to stop the
 * load (which then returns TP_E_ARG and leaves no field set). */
typedef int (*tp_plane_reader)(void *user, int slice, int comp, long long s0, long long s1,
                               float *dst);

/* Loads one slice (two_slices = 0, as tp_gpu_set_field_grid) or two
 * (two_slices = 1, as tp_gpu_set_field_grid2 with grid_dt) from read.
 * stage_bytes is the size of each pinned staging buffer (0: 32 MiB; at least
 * one slab of the six components is staged at a time).  The context keeps
 * its two staging buffers (pinned host memory and device memory of that
 * size each) for later loads, until tp_gpu_finalize. */
int tp_gpu_load_field_grid(tp_ctx *ctx, int nx, int ny, int nz, const double lo[3],
                           const double hi[3], int order, int two_slices, double e0, double b0,
                           double grid_dt, int gather_mode, tp_plane_reader read, void *user,
                           size_t stage_bytes);
/* The same from host arrays: planes_next (and planes_prev, or NULL for one
 * slice) hold the six components' planes one after another, 6*nx*ny*nz
 * floats in the given order, unscaled. */
int tp_gpu_set_field_grid_planes(tp_ctx *ctx, int nx, int ny, int nz, const double lo[3],
                                 const double hi[3], int order, double e0, double b0,
                                 const float *planes_next, const float *planes_prev,
                                 double grid_dt, int gather_mode);
/* Copies the nodes of a grid field back to the host: slice 0 (the next or
 * only slice) or 1 (the previous slice) as 8 floats per node, x fastest:
 * ex ey ez bx by bz 0 0 (the layout the gathers read).  nx, ny, nz must be
 * the grid's (TP_E_ARG otherwise); TP_E_STATE without a grid field. */
int tp_gpu_get_field_nodes(tp_ctx *ctx, int slice, int nx, int ny, int nz, float *nodes);
/* system 1: nwlcs wire-loop current systems; rows of 14 doubles in the
 * column order of the WLCS configuration file read by read_wlcs
 * (src/wlcs.c:97-179): cur_wr x_wr y_wr z_wr theta phi t0 cur_lp r_lp
 * x_lp y_lp z_lp alpha beta; plus iBessel, omega and normI
 * (src/wlcs.c:181-212).  mode TP_GATHER_REF_COMPAT evaluates the complete
 * elliptic integrals as the reference's ellint does (src/wlcs.c:460-809,
 * which is wrong for k^2 >= 0.9); TP_GATHER_TEXTBOOK evaluates them
 * exactly. */
int tp_gpu_set_field_wlcs(tp_ctx *ctx, int nwlcs, const double *rows,
                          int ibessel, double omega, double norm_i, int mode);
/* The same systems from the table read_wlcs keeps in memory: rows of 19
 * doubles in the member order of the reference's struct wlcs
 * (include/wlcs.h:24-37: x_wr y_wr z_wr costheta_wr cosphi_wr sintheta_wr
 * sinphi_wr cur_wr t0 x_lp y_lp z_lp r_lp cosalpha_lp cosbeta_lp sinalpha_lp
 * sinbeta_lp cur_lp omega), currents already scaled by normI. */
int tp_gpu_set_field_wlcs_table(tp_ctx *ctx, int nwlcs, const double *table,
                                int ibessel, int mode);
/* multi-mode turbulence (no counterpart in the reference): B = b0z z-hat +
 * sum_n a_n cos(k_n . x + phi_n), E = 0; rows of 7 doubles
 * (kx, ky, kz, ax, ay, az, phi). */
int tp_gpu_set_field_modes(tp_ctx *ctx, int nmodes, const double *rows,
                           double b0z);

/* --- fixed-step tracking (particle_tracking_fixed_time_step,
 *     src/tracking.c:90-190) ---------------------------------------------- */
/* One call does what the reference's loop does for traj_diagnose = 0:
 *   ptl[n]        in/out particles (reference: ptl)
 *   ptl_init[n]   initial particles for the diffusion terms (ptl_init)
 *   nsteps[n]     out: steps each particle was pushed, starting from 1
 *                 (nsteps_ptl_tracking, src/tracking.c:127-129)
 *   espect        out: [nt_out][nbins] particle counts inside the box
 *   espect_escape out: [nt_out][nbins] counts of escaping particles (bc 1)
 *   dcoeff        out: [8][nd] sums drr dxx dyy dzz dpp duu daa dee with
 *                 nd = (estep - sstep + 1)/dcoeff_interval + 1 rows for
 *                 sstep = 1; row k holds step k*dcoeff_interval, row 0 is 0
 *   nptl_remain   out: [nd] remaining particles (nptl_remain_time)
 * Any of espect/espect_escape/dcoeff/nptl_remain/nsteps may be NULL. */
int tp_gpu_track_fixed(tp_ctx *ctx, const tp_track_params *p, int n,
                       tp_particle *ptl, const tp_particle *ptl_init,
                       int *nsteps, double *espect, double *espect_escape,
                       double *dcoeff, int *nptl_remain);

/* The same with trajectory diagnostics, the reference's traj_diagnose = 1
 * call (src/diagnostics.c:545-557; recording at src/tracking.c:440-447):
 * particle i is copied into ptl_time[ntraj_accum[i-1] + k], k = 1, 2, ...
 * (ntraj_accum[-1] read as 0), after every step j with
 * j % nsteps_output == 0 at which it is pushed.  Slot ntraj_accum[i-1] + 0
 * holds the caller's initial point, as the reference fills it; slots that
 * are not written keep the caller's contents.  points[i] (nullable) gets
 * the final k; points past the particle's ntraj_accum[i] - ntraj_accum[i-1]
 * slots are counted but not written (the reference writes past them).
 * Not available with the surrogate field. */
int tp_gpu_track_fixed_traj(tp_ctx *ctx, const tp_track_params *p, int n,
                            tp_particle *ptl, const tp_particle *ptl_init, int *nsteps,
                            double *espect, double *espect_escape, double *dcoeff,
                            int *nptl_remain, int nsteps_output, const int *ntraj_accum,
                            tp_particle *ptl_time, int *points);

/* The same split into steps, for callers that keep particles resident in
 * HBM across calls (and for timing the tracking alone): */
int tp_gpu_load_particles(tp_ctx *ctx, int n, const tp_particle *ptl,
                          const tp_particle *ptl_init);
int tp_gpu_run_fixed(tp_ctx *ctx, const tp_track_params *p);
int tp_gpu_store_particles(tp_ctx *ctx, tp_particle *ptl, int *nsteps);
int tp_gpu_read_diagnostics(tp_ctx *ctx, double *espect,
                            double *espect_escape, double *dcoeff,
                            int *nptl_remain);
/* The same run as tp_gpu_run_fixed with no device or host allocation, for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The outputs are DEVICE buffers the caller owns, all required:
 *   espect, espect_escape  [nt_out][nbins] particle counts (unsigned)
 *   dcoeff                 [8][nd] D sums, as tp_gpu_track_fixed
 *   nptl_remain            [nd] remaining particles
 * Every buffer is cleared by the call (with kernels, no memset).  The
 * library's own scratch (D partial-sum rows, open-boundary index lists, the
 * non-finite flag) is a static device array of about 36 MB that every GPU
 * the library runs on holds until the process ends; it bounds the run to
 * TP_NOALLOC_MAX_PARTICLES resident particles (TP_E_ARG beyond).  Calls on
 * one GPU share that array, so calls that overlap on one GPU (from two
 * contexts or threads) race.  Not available with the surrogate field or
 * trajectory recording (TP_E_ARG).  Returns when the run has finished;
 * tp_gpu_last_kernel_time gives its device time as one launch span. */
#define TP_NOALLOC_MAX_PARTICLES 4194304
int tp_gpu_run_fixed_noalloc(tp_ctx *ctx, const tp_track_params *p, unsigned *espect,
                             unsigned *espect_escape, double *dcoeff, int *nptl_remain);
/* Time of the tracking kernels of the last run, measured with HIP events on
 * the stream they run on: total milliseconds and number of launches. */
int tp_gpu_last_kernel_time(tp_ctx *ctx, double *ms, int *launches);

/* --- adaptive Bulirsch-Stoer tracking ------------------------------------ */
/* The reference's adaptive stepper (src/unused/StepperBS.c:103-435, NR3
 * ch. 17 without dense output: nseq = 2(k+1), polynomial extrapolation,
 * controller constants STEPFAC1..4 = 0.65/0.94/0.02/4.0, KFAC1/2 = 0.8/0.9)
 * integrates y = (x, u), u = gamma v, with
 *   dx/dt = u/gamma,  du/dt = qm (E + (u/gamma) x B)
 * (the reference's derivs, :443-471, is not this ODE) and keeps its
 * controller state per particle (the reference keeps it in statics,
 * :167-168).  Particles run in a persistent kernel that refills lanes from
 * a work queue as particles finish. */
typedef struct tp_bs_params {
    double rtol, atol;      /* error control; the reference: 1e-14, 1e-5 */
    double qm;              /* charge/mass */
    double t_end;           /* end time, used when t_end_gyro <= 0 */
    double t_end_gyro;      /* end time per particle in its gyroperiods
                               T_g = 2 pi gamma0 / (|qm| b_ref), from its t */
    double b_ref;           /* field scale of T_g */
    double h0;              /* first trial step; <= 0: 0.01 T_g */
    int kmax;               /* extrapolation depth KMAXX, 3..8 (reference 8) */
    int max_attempts;       /* cap on step attempts per particle (0: 10^7);
                               a particle reaching it is counted as failed */
    int block;              /* lanes per block, 64/128/256 (0: tuned or 64) */
    int waves_per_cu;       /* resident waves per compute unit (0: tuned or 8) */
    int use_graph;          /* launch through a captured hipGraph */
} tp_bs_params;

typedef struct tp_bs_stats {
    long long accepted, rejected, rhs_evals;
    int failed;             /* particles stopped at max_attempts */
    int launches;
    double kernel_ms;       /* device time of the work-queue kernel */
} tp_bs_stats;

/* Resident particles (tp_gpu_load_particles) are advanced to their end
 * times; tp_gpu_store_particles returns them, its nsteps = accepted steps. */
int tp_gpu_run_adaptive(tp_ctx *ctx, const tp_bs_params *p, tp_bs_stats *stats);
/* The adaptive loop of the reference's tracking_method 1 (the commented
 * particle_tracking_adaptive, src/tracking.c:518-630) with the fixed-step
 * loop's diagnostics.  Resident particles (tp_gpu_load_particles) are
 * integrated from their time t0 to t0 + estep*dt of q (sstep must be 1,
 * precision TP_FP64; t_end and t_end_gyro of p are not used) by the
 * stepper above with NR3's dense output (src/unused/StepperBS.c:478-644:
 * nseq = 4k+2, a step is redone when its interpolation error exceeds 10
 * and the next one is capped by it), particle_bc (src/tracking.c:
 * 1172-1221) after every accepted step.  At the output times of the
 * fixed-step loop, t0 + j*dt, the state comes from the dense-output
 * polynomial of the step that contains the time (the final state at the
 * end time): the spectrum at j % einterval_t == 0, the D terms at j %
 * dcoeff_interval == 0.  An escaping particle (bc 1) is binned into the
 * escape spectrum at the frame of the next output time and stops; its
 * later D rows carry its stale terms, and the remaining counts count it as
 * tp_gpu_run_fixed does.  D terms are kept per particle and row (8 * nd *
 * n doubles, TP_E_NOMEM if that and the dense scratch exceed half the free
 * memory) and summed in a fixed order.  tp_gpu_read_diagnostics returns
 * the results, tp_gpu_store_particles the particles (nsteps = accepted
 * steps). */
int tp_gpu_run_adaptive_diag(tp_ctx *ctx, const tp_bs_params *p, const tp_track_params *q,
                             tp_bs_stats *stats);
/* The adaptive loop with the diagnostics above and the trajectory records
 * of tp_gpu_track_fixed_traj, the reference's tracking_method = 1 call with
 * traj_diagnose = 1 (src/diagnostics.c:545-557, the adaptive loop at
 * src/tracking.c:518-630, recording at src/tracking.c:440-447): load, run,
 * store and read back in one call.  Particle i's point k is its state at
 * the output time t0 + k*nsteps_output*dt <= t0 + estep*dt, from the
 * dense-output polynomial of the accepted step that contains the time (the
 * final state at the end time), folded into the box with its shifts as
 * particle_bc folds the fixed-step records (bc 0); an escaping particle
 * records none after its escape.  Slots, points and the surrogate field
 * as tp_gpu_track_fixed_traj; nsteps = accepted steps. */
int tp_gpu_track_adaptive_diag_traj(tp_ctx *ctx, const tp_bs_params *bs,
                                    const tp_track_params *p, int n, tp_particle *ptl,
                                    const tp_particle *ptl_init, int *nsteps, double *espect,
                                    double *espect_escape, double *dcoeff, int *nptl_remain,
                                    int nsteps_output, const int *ntraj_accum,
                                    tp_particle *ptl_time, int *points, tp_bs_stats *stats);
/* load + run + store in one call (nsteps may be NULL) */
int tp_gpu_track_adaptive(tp_ctx *ctx, const tp_bs_params *p, int n, tp_particle *ptl,
                          int *nsteps, tp_bs_stats *stats);
/* Launch-shape autotuner: times the work-queue kernel on the first (up to
 * 65,536) resident particles over a short horizon for each candidate
 * (block, waves_per_cu), keeps the state unchanged, writes the choice to
 * *block and *waves_per_cu and, if cache_path is given, caches it there
 * keyed by field kind, mode count and tolerance (read back on later calls). */
int tp_gpu_tune_adaptive(tp_ctx *ctx, const tp_bs_params *p, const char *cache_path,
                         int *block, int *waves_per_cu);

/* --- field probe -------------------------------------------------------- */
/* get_emf (src/emfields.c:32-60) at n points pts[4*i] = (x, y, z, t);
 * out[6*i] = (Bx, By, Bz, Ex, Ey, Ez).  precision as in tp_track_params. */
int tp_gpu_eval_field(tp_ctx *ctx, int n, const double *pts, double *out,
                      int precision);

/* --- Poincare map of the magnetic field ------------------------------------
 * run_type 2 of the reference (src/poincare_map.c:39-258): each field line is
 * traced from its start point by classical RK4 on the unit vector B/|B| with
 * the field at t = 0 (next_location, :268-320), every stage point folded once
 * into the periodic box (field_boundary, :326-352), and the positions where it
 * crosses the plane face = start[face] are recorded (:166-184).  The field
 * is the context's field; the reference's systems 3 and 4 are the PIC grid
 * kinds (tp_gpu_set_field_grid, _grid2 evaluated at t = 0); the analytic
 * kinds have no counterpart there. */
#define TP_PMAP_REF_COMPAT 0  /* the crossing test of src/poincare_map.c:168-173:
                                 the face coordinate before a step is read before
                                 that step's periodic fold, so a fold of the face
                                 axis counts as a crossing, and the first step
                                 from the start counts as one */
#define TP_PMAP_PERIODIC   1  /* a crossing is a sign change within one step of
                                 the distance to the plane reduced into
                                 [-L/2, L/2] (L the box length along face) */

typedef struct tp_pmap_params {
    int face;               /* normal of the cross section: 0 x, 1 y, 2 z */
    int ncross;             /* crossings recorded per field line, >= 0 */
    int mode;               /* TP_PMAP_* */
    int chunk;              /* RK4 steps per line per kernel launch (0: 16384) */
    int precision;          /* TP_FP64: the reference's arithmetic, bitwise;
                               TP_FP32_MIXED: field, unit vector and RK4 sum in
                               FP32, positions, traced length and crossing test
                               in FP64 */
    double step;            /* field-line step, > 0 */
    double length_max;      /* traced length at which a line stops */
    double lo[3], hi[3];    /* periodic box (simul_domain) */
    int lanes;              /* lanes per field line: 0 automatic (one; four,
                               one per component, for the launches over at
                               most TP_PMAP_QUAD_LINES running lines, on
                               reference-gather PIC grids in TP_FP64), 1 always
                               one, 4 four whenever the field allows; the
                               records are the same bits either way */
} tp_pmap_params;
#define TP_PMAP_QUAD_LINES 32768

/* Traces npoints field lines from starts[3*i] (host arrays).  xpos, ypos,
 * zpos have npoints*(ncross+1) entries each (the reference's datasets,
 * src/poincare_map.c:196-240): slot i*(ncross+1) is start i, slot
 * i*(ncross+1)+k its k-th crossing (the position after the step that
 * crossed); slots not reached keep 2*hi (:55-80).  ncrossed[i] and
 * nsteps[i] (either may be NULL) receive the crossings found and the RK4
 * steps taken. */
int tp_gpu_poincare_map(tp_ctx *ctx, const tp_pmap_params *p, int npoints,
                        const double *starts, double *xpos, double *ypos, double *zpos,
                        int *ncrossed, long long *nsteps);

/* --- learned field surrogate ---------------------------------------------
 * No counterpart in the reference (its closest slot is the field dispatch
 * get_emf, src/emfields.c:32-60).  A coordinate network stands in for the
 * multi-mode field of tp_gpu_set_field_modes:
 *   f   = [sin 2 pi (q_j . x), cos 2 pi (q_j . x)]   j < 128 (256 features)
 *   h1  = gelu(f W1 + b1), h2 = gelu(h1 W2 + b2), h3 = gelu(h2 W3 + b3)  (128 each)
 *   B   = b0z z-hat + scale (h3 W4 + b4)
 * with gelu the tanh form z / (1 + exp(-2 sqrt(2/pi) (z + 0.044715 z^3)))
 * as the kernels evaluate it: max(z, 0) minus a degree-8 polynomial for its
 * tail, in packed FP16 (within 2^-10 |z| + 1.1e-3 of the tanh form; its
 * derivative in training is the tanh form's).  Layers
 * 1-3 run on the matrix cores in bfloat16 with FP32 accumulation; layer 4,
 * the biases and the frequencies are FP32.  The frequencies q_j = n_j / L
 * with integer n_j make the network periodic in the box L.
 * params: TP_MLP_NPARAMS floats, row index = input:
 *   W1[256][128] b1[128] W2[128][128] b2[128] W3[128][128] b3[128] W4[128][3] b4[3] */
#define TP_MLP_NFREQ    128
#define TP_MLP_NPARAMS  66307

typedef struct tp_mlp_train_params {
    long long nsamples;   /* training set size (points drawn uniformly in [0, period)^3) */
    int batch;            /* samples per optimiser step, a multiple of 256 */
    int epochs;           /* passes over the training set */
    double period;        /* L, the box of the samples */
    double lr, lr_min;    /* Adam step size, cosine-annealed from lr to lr_min */
    double beta1, beta2, eps;
    unsigned long long seed;  /* sample positions: splitmix64 of (seed, index) */
} tp_mlp_train_params;

typedef struct tp_mlp_train_stats {
    double first_loss;    /* mean loss of the first epoch's first step */
    double last_loss;     /* mean loss over the last epoch */
    long long steps;      /* optimiser steps taken */
    double sample_ms;     /* time to generate the training set (exact FP64 field) */
    double train_ms;      /* time of the optimiser steps */
} tp_mlp_train_stats;

/* Sets the network (host arrays: q[3*128] cycles per unit length,
 * params[TP_MLP_NPARAMS]) and clears the optimiser state. */
int tp_gpu_mlp_set(tp_ctx *ctx, const float *q, const float *params, double scale);
/* Copies the current FP32 parameters out. */
int tp_gpu_mlp_get(tp_ctx *ctx, float *params);
/* Network outputs (before scale and guide field) at n points x[3*i]. */
int tp_gpu_mlp_forward(tp_ctx *ctx, int n, const double *x, double *out);
/* Mean-squared-error loss of the outputs against y[3*i] over n samples and
 * its gradient with respect to the parameters (no update). */
int tp_gpu_mlp_loss_grad(tp_ctx *ctx, int n, const float *x, const float *y, float *grad,
                         double *loss);
/* Trains the network on samples of the context's mode field (set by
 * tp_gpu_set_field_modes; targets are the exact FP64 mode sums divided by
 * scale, without the guide field) with the kernels above and Adam. */
int tp_gpu_train_surrogate(tp_ctx *ctx, const tp_mlp_train_params *p,
                           tp_mlp_train_stats *stats);
/* Makes the network the field source of tp_gpu_run_fixed (FP32 Boris only),
 * with the context's modes as the exact field.  Before the run's first
 * launch, and before every launch that starts after a multiple of
 * check_interval steps (0: the first only), each particle still on the
 * network compares it with the exact FP32 mode sum at its position; once
 * |B_net - B_exact| / |B_exact| exceeds tol it uses the exact field for the
 * rest of the run.  Network and exact-field particles are run by separate
 * kernels (no wave mixes them).  The D steps' pitch angles use the exact
 * FP64 field. */
int tp_gpu_set_field_mlp(tp_ctx *ctx, double tol, int check_interval);
/* Particles that switched to the exact field during the last run. */
int tp_gpu_mlp_fallback_count(tp_ctx *ctx, long long *count);
/* Per resident particle (n = the count loaded last): 1 if it used the exact
 * field at the end of the last surrogate run, else 0. */
int tp_gpu_mlp_fallback_flags(tp_ctx *ctx, int n, int *flags);

#ifdef __cplusplus
}
#endif
#endif /* TP_GPU_H */
