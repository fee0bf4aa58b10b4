/*
 * tp_track.h -- step-kernel arguments and the per-step helpers shared by the
 * fixed-step kernels (tp_gpu.hip) and the surrogate-field kernel (tp_mlp.hip).
 */
#ifndef TP_TRACK_H
#define TP_TRACK_H

#include <hip/hip_runtime.h>
#include "tp_device.h"

/* ptl_energy_fixed binning, src/diagnostics.c:193-203 */
static __device__ __noinline__ int energy_bin(double ux, double uy, double uz,
                                          int nbins, double logemin, double logde)
{
    double gama = sqrt(1.0 + ux * ux + uy * uy + uz * uz);
    double ene = gama - 1.0;
    double v = floor((log10(ene) - logemin) / logde);
    if (!(v > -2147483649.0 && v < 2147483648.0)) return 0;
    int ib = (int)v;
    if (ib < 0) ib = 0;
    else if (ib > nbins - 1) ib = nbins - 1;
    return ib;
}

struct TrackArgs {
    FieldDev f;
    int n, bc, nbins, nt_out, einterval, di, fix_stale, first;
    int j0, j1, ntp0, nd;
    int dstep;           /* step kernels: D row evaluated after step j1 (-1: none) */
    int ntot;            /* all particles of the run: the stride of st */
    int pboff;           /* D partial sums go to block rows pboff + blockIdx.x */
    double dt, qm, lo[3], hi[3], logemin, logde;
    double *x, *y, *z, *ux, *uy, *uz, *t, *xs, *ys, *zs;
    const double *x0, *y0, *z0, *ux0, *uy0, *uz0, *g0, *mu0, *th0;
    double *st;
    int *esc, *nsteps;
    unsigned *espect, *espect_esc;
    double *partials;
    int *rdec;
    const int *idx;      /* step kernels: particle g of the launch is idx[g] (null: g) */
    /* trajectory diagnostics (traj != null): records of particle i go to
     * traj[traj_accum[i-1] + k], k = traj_pts[i], ... every traj_nout steps;
     * qm of the records from stage[i] (the particles as uploaded) */
    tp_particle *traj;
    const int *traj_accum;
    int *traj_pts;
    int traj_nout;
    const tp_particle *stage;
};

/* particle_bc, src/tracking.c:1172-1221 */
static __device__ __forceinline__ void particle_bc(const TrackArgs &a, Ptl &p, double &xs,
                                            double &ys, double &zs, int &esc)
{
    if (a.bc == 0) {
        /* a wave none of whose particles left the box skips the wraps */
        if (p.x > a.hi[0] || p.y > a.hi[1] || p.z > a.hi[2] || p.x < a.lo[0] ||
            p.y < a.lo[1] || p.z < a.lo[2]) {
            double xd = a.hi[0] - a.lo[0], yd = a.hi[1] - a.lo[1], zd = a.hi[2] - a.lo[2];
            if (p.x > a.hi[0]) { p.x -= xd; xs += xd; }
            if (p.y > a.hi[1]) { p.y -= yd; ys += yd; }
            if (p.z > a.hi[2]) { p.z -= zd; zs += zd; }
            if (p.x < a.lo[0]) { p.x += xd; xs -= xd; }
            if (p.y < a.lo[1]) { p.y += yd; ys -= yd; }
            if (p.z < a.lo[2]) { p.z += zd; zs -= zd; }
        }
    } else if (a.bc == 1) {
        if (p.x > a.hi[0] || p.y > a.hi[1] || p.z > a.hi[2] ||
            p.x < a.lo[0] || p.y < a.lo[1] || p.z < a.lo[2])
            esc = 1;
    } else if (a.bc == 2) {
        esc = 0;
    }
}

/* the surrogate field (tp_mlp.hip): run set-up and one launch (residual
 * check and split into network and exact-field particles when `check`,
 * then the two step kernels; exact_fn launches the exact-field kernel) */
struct tp_ctx;
typedef void (*track_fn)(dim3, dim3, size_t, hipStream_t, const TrackArgs &);
int tp_mlp_prepare_run(tp_ctx *c, int n);
int tp_mlp_launch(tp_ctx *c, const TrackArgs &a, size_t shmem, int check, int block,
                  track_fn exact_fn);

#endif /* TP_TRACK_H */
