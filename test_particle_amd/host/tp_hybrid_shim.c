/*
 * tp_hybrid_shim.c -- GPU drop-in for the reference tracer's tracking loop.
 *
 * Build this file INTO THE REFERENCE PROGRAM (it includes the reference's
 * headers and reads its globals) and link libtpgpu.so.  It defines
 * particle_tracking_hybrid with exactly the reference's signature
 * (src/tracking.c:659-663, include/tracking.h:78-82); compile the reference's
 * src/tracking.c with -Dparticle_tracking_hybrid=particle_tracking_hybrid_cpu
 * so the CPU loop stays available for what the GPU path does not cover
 * (see tp_shim_set_field; systems 2 and 3 with several time frames, which
 * the reference's loop does not track either).  Trajectory runs
 * (traj_diagnose = 1) go to tp_gpu_track_fixed_traj (tracking_method 0) or
 * tp_gpu_track_adaptive_diag_traj (tracking_method 1), two PIC time slices
 * to tp_gpu_set_field_grid2, system 3 to tp_gpu_set_field_mhd.
 *
 * The field source is taken from the reference's globals once per call;
 * results go through the reference's own collect_espectrum and
 * collect_diff_coeffs (src/diagnostics.c:221-248, 723-782), so the MPI
 * reductions and output files are unchanged.
 *
 * Which path ran is visible on stderr: rank 0 prints
 *   tp_gpu: particle_tracking_hybrid on the GPU (...)
 * when the GPU tracks, and every rank that runs the CPU loop prints
 *   tp_gpu: particle_tracking_hybrid runs the CPU loop: <reason>
 * With TP_SHIM_REQUIRE_GPU=1 in the environment a call the GPU path does not
 * take aborts the program (MPI_Abort) instead of running the CPU loop.
 *
 * The optional "GPU tracking" section of init.dat (tp_shim_config.h) picks
 * the pusher (Wirz or Boris), the precision (FP64 or the FP32 form), the
 * non-finite check, the stale-D fix and the adaptive stepper's tolerances,
 * first step and depth; it is read once, at the first call.  A malformed
 * section, or a setting other than the reference's Wirz/FP64 loop when the
 * CPU loop would have to run, aborts the program.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include <mpi.h>
#include "constants.h"
#include "domain.h"
#include "particle_info.h"
#include "tracking.h"
#include "diagnostics.h"
#include "force_free.h"
#include "magnetic_field.h"
#include "electric_field.h"
#include "velocity_field.h"
#include "wlcs.h"
#include "tp_gpu.h"
#include "tp_shim_config.h"

void particle_tracking_hybrid_cpu(int mpi_rank, int nptl, double dt, int nbins,
        int nt_out, int bc_flag, int nsteps_output, double pmass,
        int *ntraj_accum, int tracking_method, int traj_diagnose,
        particles *ptl_init, int *nsteps_ptl_tracking, particles *ptl,
        particles *ptl_time);

/* reference globals (src/tracking.c:35-40, src/force_free.c:33,
 * src/{magnetic,electric,velocity}_field.c:34-39, src/wlcs.c:12-14,
 * src/diagnostics.c:32) */
extern struct domain simul_domain;
extern struct grids simul_grid;
extern int system_type, multi_tframe, data_type;
extern double charge_mass;
extern struct param_ff param;
extern struct bfields_float *magc_float, *maga_float, *magb_float;
extern struct efields_float *efc_float, *efa_float, *efb_float;
extern struct vfields_double *vc_double;
extern struct vfields_float *vc_float;
extern struct efields_double *efc_double;
extern struct bfields_double *magc_double;
extern struct wlcs *config;
extern int nwlcs, iBessel;
extern double emin, logemin, logde;

static tp_ctx *tp_shim_ctx;
static int tp_shim_init_rc = 1;   /* 1: not tried yet */
static tp_shim_config tp_shim_cfg;
static int tp_shim_cfg_read;

/* the GPU tracking section of init.dat (src/main.c:30 reads the same file),
 * once; a malformed section stops the program as the reference's parser
 * does (src/domain.c:93-96) */
static void tp_shim_read_config(void)
{
    char err[256];
    if (tp_shim_cfg_read) return;
    tp_shim_cfg_read = 1;
    if (tp_shim_config_read("init.dat", &tp_shim_cfg, err, sizeof(err)) != 0) {
        fprintf(stderr, "tp_gpu: %s\n", err);
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
}

static const char *tp_shim_pusher_name(int pusher)
{
    return pusher == TP_PUSH_BORIS ? "boris" : "wirz";
}

/* NULL on success; otherwise why the GPU path does not take this setup */
static const char *tp_shim_set_field(tp_ctx *c)
{
    double lo[3] = {simul_domain.xmin, simul_domain.ymin, simul_domain.zmin};
    double hi[3] = {simul_domain.xmax, simul_domain.ymax, simul_domain.zmax};
    int rc;
    switch (system_type) {
    case 0:
        rc = tp_gpu_set_field_test(c);
        break;
    case 1:
        rc = tp_gpu_set_field_wlcs_table(c, nwlcs, (const double *)config, iBessel,
                                         TP_GATHER_REF_COMPAT);
        break;
    case 2:
        /* with several velocity frames the reference's fixed-step loop tracks
         * nothing (src/tracking.c:151-157); its CPU loop reproduces that */
        if (multi_tframe == 1) return "system 2 with multi_tframe = 1 is not tracked";
        rc = tp_gpu_set_velocity_grid(c, simul_grid.nx, simul_grid.ny, simul_grid.nz, lo, hi,
                                      (const double *)vc_double);
        if (rc == TP_OK) rc = tp_gpu_set_field_abc(c, ffA, ffB, ffC, param.lambda0, param.B0);
        break;
    case 3:
        /* MHD velocity and magnetic grids (getemf_mhd_test_particle,
         * src/emfields.c:107-120); with several frames the reference's loop
         * tracks nothing (src/tracking.c:151-157) */
        if (multi_tframe == 1) return "system 3 with multi_tframe = 1 is not tracked";
        if (data_type != (int)sizeof(float)) return "system 3 with double grids";
        rc = tp_gpu_set_field_mhd(c, simul_grid.nx, simul_grid.ny, simul_grid.nz, lo, hi,
                                  (const float *)vc_float, (const float *)magc_float);
        break;
    case 4:
        if (data_type == (int)sizeof(double)) {
            /* the reference's double gathers (src/electric_field.c:370-406,
             * src/magnetic_field.c:555-591); its own get_emf would read the
             * unallocated float arrays instead */
            if (multi_tframe == 1) return "two time slices of double PIC grids";
            rc = tp_gpu_set_field_grid_double(c, simul_grid.nx, simul_grid.ny, simul_grid.nz,
                                              lo, hi, (const double *)efc_double,
                                              (const double *)magc_double);
            break;
        }
        if (data_type != (int)sizeof(float)) return "PIC grids of an unknown data type";
        if (multi_tframe == 1)
            /* two time slices: efa/maga = next, efb/magb = previous
             * (src/magnetic_field.c:470-524) */
            rc = tp_gpu_set_field_grid2(c, simul_grid.nx, simul_grid.ny, simul_grid.nz, lo, hi,
                                        (const float *)efa_float, (const float *)maga_float,
                                        (const float *)efb_float, (const float *)magb_float,
                                        simul_grid.dt, TP_GATHER_REF_COMPAT);
        else
            rc = tp_gpu_set_field_grid(c, simul_grid.nx, simul_grid.ny, simul_grid.nz, lo, hi,
                                       (const float *)efc_float, (const float *)magc_float,
                                       TP_GATHER_REF_COMPAT);
        break;
    default:
        return "unknown system type";
    }
    return rc == TP_OK ? NULL : tp_gpu_strerror(rc);
}

static int tp_shim_require_gpu(void)
{
    const char *e = getenv("TP_SHIM_REQUIRE_GPU");
    return e && e[0] && strcmp(e, "0") != 0;
}

void particle_tracking_hybrid(int mpi_rank, int nptl, double dt, int nbins,
        int nt_out, int bc_flag, int nsteps_output, double pmass,
        int *ntraj_accum, int tracking_method, int traj_diagnose,
        particles *ptl_init, int *nsteps_ptl_tracking, particles *ptl,
        particles *ptl_time)
{
    int rc;
    const char *why = NULL;
    tp_shim_read_config();
    const tp_shim_config *cfg = &tp_shim_cfg;
    if (tp_shim_init_rc == 1) tp_shim_init_rc = tp_gpu_init(0, &tp_shim_ctx);
    if (tp_shim_init_rc != TP_OK)
        why = tp_gpu_strerror(tp_shim_init_rc);
    else if (tracking_method != 0 && tracking_method != 1)
        why = "tracking_method is neither 0 nor 1";
    else if (traj_diagnose != 0 && traj_diagnose != 1)
        why = "traj_diagnose is neither 0 nor 1";
    else
        why = tp_shim_set_field(tp_shim_ctx);
    if (why) {
        fprintf(stderr, "tp_gpu: particle_tracking_hybrid runs the CPU loop: %s\n", why);
        if (tp_shim_require_gpu()) {
            fprintf(stderr, "tp_gpu: TP_SHIM_REQUIRE_GPU is set; aborting\n");
            MPI_Abort(MPI_COMM_WORLD, 1);
        }
        /* the reference's loop is the Wirz pusher in FP64 with its stale D
         * terms (src/tracking.c:429); it cannot honour other settings */
        if (tracking_method == 0 && (cfg->pusher != TP_PUSH_WIRZ || cfg->precision != TP_FP64 ||
                                     (cfg->flags & TP_FLAG_FIX_STALE_DCOEFF))) {
            fprintf(stderr, "tp_gpu: the GPU tracking section of init.dat asks for what the "
                    "CPU loop does not do (pusher %s, precision %s, stale-D fix %d); "
                    "aborting\n", tp_shim_pusher_name(cfg->pusher),
                    cfg->precision == TP_FP64 ? "fp64" : "fp32",
                    (cfg->flags & TP_FLAG_FIX_STALE_DCOEFF) ? 1 : 0);
            MPI_Abort(MPI_COMM_WORLD, 1);
        }
        particle_tracking_hybrid_cpu(mpi_rank, nptl, dt, nbins, nt_out, bc_flag,
                nsteps_output, pmass, ntraj_accum, tracking_method, traj_diagnose,
                ptl_init, nsteps_ptl_tracking, ptl, ptl_time);
        return;
    }
    if (mpi_rank == 0) {
        char dev[128] = "";
        tp_gpu_device_info(tp_shim_ctx, dev, (int)sizeof(dev), NULL);
        if (tracking_method == 1)
            fprintf(stderr, "tp_gpu: particle_tracking_hybrid on the GPU (%s; system %d, "
                    "%d particles, adaptive, rtol %g, atol %g, first step %g, kmax %d)\n",
                    dev, system_type, nptl, cfg->bs_rtol, cfg->bs_atol, cfg->bs_h0,
                    cfg->bs_kmax);
        else
            fprintf(stderr, "tp_gpu: particle_tracking_hybrid on the GPU (%s; system %d, "
                    "%d particles, traj_diagnose %d, %s, %s, flags %d)\n", dev, system_type,
                    nptl, traj_diagnose, tp_shim_pusher_name(cfg->pusher),
                    cfg->precision == TP_FP64 ? "fp64" : "fp32", cfg->flags);
    }
    tp_track_params p;
    p.pusher = cfg->pusher;
    p.precision = cfg->precision;
    p.bc_flag = bc_flag;
    p.nbins = nbins;
    p.nt_out = nt_out;
    p.sstep = 1;
    p.estep = simul_domain.tmax / dt;              /* src/tracking.c:103 */
    p.einterval_t = p.estep / nt_out;              /* src/tracking.c:104 */
    p.dcoeff_interval = 100;                       /* src/tracking.c:34 */
    p.flags = cfg->flags;
    p.dt = dt;
    p.qm = charge_mass;
    p.xmin = simul_domain.xmin; p.xmax = simul_domain.xmax;
    p.ymin = simul_domain.ymin; p.ymax = simul_domain.ymax;
    p.zmin = simul_domain.zmin; p.zmax = simul_domain.zmax;
    p.logemin = logemin;
    p.logde = logde;
    int nd = p.estep / 100 + 1;                    /* src/tracking.c:109 */
    double *espect = calloc((size_t)nt_out * nbins, sizeof(double));
    double *espect_esc = calloc((size_t)nt_out * nbins, sizeof(double));
    double *tot = calloc((size_t)nt_out * nbins, sizeof(double));
    double *d = calloc((size_t)8 * nd, sizeof(double));
    int *remain = calloc(nd, sizeof(int));
    double t1 = omp_get_wtime();
    if (tracking_method == 1) {
        /* the adaptive loop (particle_tracking_adaptive, commented out in
         * src/tracking.c:518-630; its branch at :680-683 is empty): the
         * Bulirsch-Stoer stepper with dense output, by default with the
         * reference's constants (KMAXX 8, atol 1e-5, rtol 1e-14,
         * include/StepperBS.h:18-22; first step 1e-5, src/tracking.c:566),
         * spectra at the fixed-step loop's output times; like the
         * reference's adaptive branch it writes no diffusion file.  It
         * integrates the equations of motion in FP64; the pusher and
         * precision settings do not apply to it. */
        tp_bs_params b;
        memset(&b, 0, sizeof(b));
        b.rtol = cfg->bs_rtol;
        b.atol = cfg->bs_atol;
        b.qm = charge_mass;
        b.b_ref = 1.0;
        b.h0 = cfg->bs_h0;
        b.kmax = cfg->bs_kmax;
        p.precision = TP_FP64;
        if (traj_diagnose == 1) {
            /* trajectory run: ptl_time filled at the output times, no
             * spectrum files (src/tracking.c:185) */
            rc = tp_gpu_track_adaptive_diag_traj(
                tp_shim_ctx, &b, &p, nptl, (tp_particle *)ptl, (const tp_particle *)ptl_init,
                nsteps_ptl_tracking, NULL, NULL, NULL, NULL, nsteps_output, ntraj_accum,
                (tp_particle *)ptl_time, NULL, NULL);
            if (rc != TP_OK) {
                fprintf(stderr, "tp_gpu_track_adaptive_diag_traj: %s\n", tp_gpu_strerror(rc));
                MPI_Abort(MPI_COMM_WORLD, 1);
            }
            double t2 = omp_get_wtime();
            if (mpi_rank == 0) printf("Elapsed time: %lf s\n", t2 - t1);
            free(espect); free(espect_esc); free(tot); free(d); free(remain);
            return;
        }
        rc = tp_gpu_load_particles(tp_shim_ctx, nptl, (const tp_particle *)ptl,
                                   (const tp_particle *)ptl_init);
        if (rc == TP_OK) rc = tp_gpu_run_adaptive_diag(tp_shim_ctx, &b, &p, NULL);
        if (rc == TP_OK)
            rc = tp_gpu_store_particles(tp_shim_ctx, (tp_particle *)ptl, nsteps_ptl_tracking);
        if (rc == TP_OK)
            rc = tp_gpu_read_diagnostics(tp_shim_ctx, espect, espect_esc, d, remain);
        if (rc != TP_OK) {
            fprintf(stderr, "tp_gpu_run_adaptive_diag: %s\n", tp_gpu_strerror(rc));
            MPI_Abort(MPI_COMM_WORLD, 1);
        }
        double t2 = omp_get_wtime();
        if (mpi_rank == 0) printf("Elapsed time: %lf s\n", t2 - t1);
        collect_espectrum(mpi_rank, nbins, nt_out, (double (*)[nbins])espect,
                          (double (*)[nbins])tot, "data/espectrum.dat");
        if (bc_flag == 1) {
            for (size_t k = 0; k < (size_t)nt_out * nbins; k++) tot[k] = 0.0;
            collect_espectrum(mpi_rank, nbins, nt_out, (double (*)[nbins])espect_esc,
                              (double (*)[nbins])tot, "data/espect_escape.dat");
        }
        free(espect); free(espect_esc); free(tot); free(d); free(remain);
        return;
    }
    if (traj_diagnose == 1) {
        /* trajectory run (src/diagnostics.c:555-557): ptl_time filled, no
         * spectrum or diffusion files (src/tracking.c:173, :688) */
        rc = tp_gpu_track_fixed_traj(tp_shim_ctx, &p, nptl, (tp_particle *)ptl,
                                     (const tp_particle *)ptl_init, nsteps_ptl_tracking,
                                     NULL, NULL, NULL, NULL, nsteps_output, ntraj_accum,
                                     (tp_particle *)ptl_time, NULL);
        if (rc != TP_OK) {
            fprintf(stderr, "tp_gpu_track_fixed_traj: %s\n", tp_gpu_strerror(rc));
            MPI_Abort(MPI_COMM_WORLD, 1);
        }
        double t2 = omp_get_wtime();
        if (mpi_rank == 0) printf("Elapsed time: %lf s\n", t2 - t1);
        free(espect); free(espect_esc); free(tot); free(d); free(remain);
        return;
    }
    rc = tp_gpu_track_fixed(tp_shim_ctx, &p, nptl, (tp_particle *)ptl,
                            (const tp_particle *)ptl_init, nsteps_ptl_tracking, espect,
                            espect_esc, d, remain);
    if (rc != TP_OK) {
        fprintf(stderr, "tp_gpu_track_fixed: %s\n", tp_gpu_strerror(rc));
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    collect_diff_coeffs(mpi_rank, nd, d, d + nd, d + 2 * nd, d + 3 * nd, d + 4 * nd,
                        d + 5 * nd, d + 6 * nd, d + 7 * nd, "data/diff_coeffs.dat", remain);
    double t2 = omp_get_wtime();
    if (mpi_rank == 0) printf("Elapsed time: %lf s\n", t2 - t1);
    collect_espectrum(mpi_rank, nbins, nt_out, (double (*)[nbins])espect,
                      (double (*)[nbins])tot, "data/espectrum.dat");
    if (bc_flag == 1) {
        for (size_t k = 0; k < (size_t)nt_out * nbins; k++) tot[k] = 0.0;
        collect_espectrum(mpi_rank, nbins, nt_out, (double (*)[nbins])espect_esc,
                          (double (*)[nbins])tot, "data/espect_escape.dat");
    }
    free(espect); free(espect_esc); free(tot); free(d); free(remain);
}
