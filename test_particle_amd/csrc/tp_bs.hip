/*
 * tp_bs.hip -- adaptive Bulirsch-Stoer tracking on gfx950.
 *
 * Algorithm: the reference's StepperBS (src/unused/StepperBS.c:103-435, NR3
 * ch. 17 without dense output), with per-particle controller state and the
 * relativistic ODE in (x, u).  The CPU oracle (tpref_bs_track) runs the same
 * operations as plain nested loops; this file reorganises them for a GPU:
 *
 *  - Every loop iteration of the kernel performs exactly one right-hand-side
 *    evaluation on every active lane.  A lane's position inside a BS step
 *    (which extrapolation row k, which midpoint substep, whether the next
 *    evaluation is the derivative at an accepted point) is a small state
 *    machine, so lanes at different rows and orders still evaluate the field
 *    together and the expensive field sum (1000 cosines for the multi-mode
 *    field) never diverges across a wavefront.  Only the cheap tableau and
 *    controller bookkeeping between evaluations is divergent.
 *  - The kernel is persistent: a fixed number of waves per compute unit pull
 *    particles from a global work queue.  Lanes whose particle reached its
 *    end time are refilled together: one atomic per wavefront hands out as
 *    many particle indices as there are idle lanes.
 *  - The counter reset and the kernel are captured once into a hipGraph and
 *    replayed (the reset is a small kernel, not a memset node).
 *  - Launch shape (lanes per block, waves per CU) comes from an autotuner
 *    that times short runs of the same kernel and caches its choice.
 *
 * The controller's pow() is tp_pow (shared with the oracle), so step-size
 * decisions and therefore trajectories are bitwise those of the oracle.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tp_ctx.h"

#define BS_N 6
#define BS_STEPFAC1 0.65
#define BS_STEPFAC2 0.94
#define BS_STEPFAC3 0.02
#define BS_STEPFAC4 4.0
#define BS_KFAC1 0.8
#define BS_KFAC2 0.9

enum { PH_INIT = 0, PH_MID = 1, PH_ACCEPT = 2, PH_DROW = 3, PH_PREP = 4 };

/* Dense form: a lane whose step was accepted waits (PH_PREP, evaluating a
 * dummy point like an idle lane) until this many lanes of its wavefront
 * wait, or no other lane is busy, and then they form their interpolants
 * together.  prepare_dense is a long chain of per-lane scratch accesses; run
 * for one lane at a time it stalled the whole wavefront at almost every
 * iteration.  Each lane's arithmetic is unchanged (same results). */
#ifndef BS_PREP_BATCH
#define BS_PREP_BATCH 16
#endif
enum { Q_HEAD = 0, Q_ACC = 1, Q_REJ = 2, Q_RHS = 3, Q_FAIL = 4, Q_N = 8 };

struct BsConsts {
    int kmax, ktarg0, dense;
    int nseq[10], cost[10], ipoint[11];
    double coeff[10][10];
    double errfac[19];
};

/* dense output: per lane, in HBM, ysave (9 x 6), fsave (162 x 6: every
 * substep derivative of sequences 0..8, ipoint[9] = 162 rows) and the
 * interpolation coefficients (23 x 6); slot s of lane g at ds[s * nl + g] */
#define DS_YSAVE 0
#define DS_FSAVE (9 * BS_N)
#define DS_DENS (DS_FSAVE + 162 * BS_N)
#define DS_YV (DS_DENS + (2 * 9 + 5) * BS_N)
/* largest kmax of the dense form; YV: the values each interpolation order
 * takes from each sequence, for one component at a time ([order][sequence]) */
#define BS_KMAX 8
#define DS_SLOTS (DS_YV + (2 * BS_KMAX - 1) * (BS_KMAX + 1))

struct BsArgs {
    FieldDev f;
    BsConsts c;
    int n, cap;
    double qm, rtol, atol, t_end, t_end_gyro, b_ref, h0;
    double *x, *y, *z, *ux, *uy, *uz, *t;
    int *nsteps, *flag;
    unsigned long long *ctr;
    /* the dense form with the fixed-step loop's diagnostics */
    double *ds;                  /* dense-output scratch, DS_SLOTS per lane */
    long nl;                     /* lanes of the launch */
    double dt;                   /* output grid: times t0 + j dt */
    int estep, ei, di, nd, nframes, nbins, nt_out, bc, fix_stale;
    double lo[3], hi[3], logemin, logde;
    double *xs, *ys, *zs;        /* periodic shifts */
    const double *x0, *y0, *z0, *ux0, *uy0, *uz0, *g0, *mu0, *th0;
    double *st;                  /* 5 stale D terms per particle (SoA) */
    unsigned *espect, *espect_esc;
    double *dbuf;                /* D terms [row][8][n] */
    int *outhist;                /* particles counted out from row r, [nd] */
    /* trajectory diagnostics (traj != null): the particle at the output
     * steps j = k traj_nout into traj[traj_accum[i-1] + traj_pts[i]++]
     * while that slot is the particle's; qm from stage[i] */
    tp_particle *traj;
    const int *traj_accum;
    int *traj_pts;
    int traj_nout;
    const tp_particle *stage;
};

/* StepperBase (src/unused/StepperBS.c:103-149): the step-size sequence
 * (dense output: nseq = 4k+2, else 2(k+1)), costs, k_targ from rtol, the
 * extrapolation coefficients, the dense interpolation error factors and
 * ipoint, where each sequence's derivatives start in fsave */
static void bs_consts(int kmax, double rtol, int dense, BsConsts *c)
{
    memset(c, 0, sizeof(*c));
    c->kmax = kmax;
    c->dense = dense;
    for (int i = 0; i <= kmax; i++) c->nseq[i] = dense ? 4 * i + 2 : 2 * (i + 1);
    c->cost[0] = c->nseq[0] + 1;
    for (int k = 0; k < kmax; k++) c->cost[k + 1] = c->cost[k] + c->nseq[k + 1];
    double logfact = -log10(rtol > 1.0e-12 ? rtol : 1.0e-12) * 0.6 + 0.5;
    int kt = (int)logfact;
    if (kt > kmax - 1) kt = kmax - 1;
    if (kt < 1) kt = 1;
    c->ktarg0 = kt;
    for (int k = 0; k <= kmax; k++)
        for (int l = 0; l < k; l++) {
            double ratio = (double)c->nseq[k] / c->nseq[l];
            c->coeff[k][l] = 1.0 / (ratio * ratio - 1.0);
        }
    for (int i = 0; i < 2 * (kmax + 1) + 1; i++) {
        int ip5 = i + 5;
        c->errfac[i] = 1.0 / (ip5 * ip5);
        double e = 0.5 * sqrt(((double)(i + 1)) / ip5);
        for (int j = 0; j <= i; j++) c->errfac[i] *= e / (j + 1);
    }
    c->ipoint[0] = 0;
    for (int i = 1; i <= kmax + 1; i++) {
        int njadd = 4 * i - 2;
        if (c->nseq[i - 1] > njadd) njadd++;
        c->ipoint[i] = c->ipoint[i - 1] + njadd;
    }
}

/* the right-hand side from the field e at the point y */
__device__ __forceinline__ void bs_deriv(const Emf &e, double qm, const double *y, double *dy)
{
    double g = sqrt(1.0 + y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
    double ig = 1.0 / g;
    double vx = y[3] * ig, vy = y[4] * ig, vz = y[5] * ig;
    dy[0] = vx;
    dy[1] = vy;
    dy[2] = vz;
    dy[3] = qm * (e.ex + (vy * e.bz - vz * e.by));
    dy[4] = qm * (e.ey + (vz * e.bx - vx * e.bz));
    dy[5] = qm * (e.ez + (vx * e.by - vy * e.bx));
}

template <int FK>
__device__ __forceinline__ void bs_rhs(const FieldDev &f, double qm, double t, const double *y,
                                       double *dy)
{
    Emf e;
    get_emf64<FK>(f, y[0], y[1], y[2], t, e);
    bs_deriv(e, qm, y, dy);
}

/* per-lane state of one particle's integration */
struct BsLane {
    long pi;                    /* particle, -1 idle, -2 retired */
    int phase, k, nn, nstep, ktarg, attempts, acc, rej;
    int first, last, reject, prev_reject;
    double t, tend, h, hnew, hnext, hsub, xe;
    /* yv: the last accepted state, which is also the start of the step being
     * attempted (the reference's ysav) */
    double yv[BS_N], ycur[BS_N], dydx[BS_N], ym[BS_N], yn[BS_N];
    double tab[8][BS_N], hopt[9], work[9];
    int rhs;
    /* dense form: next fsave row, the polynomial's step (start, size), the
     * step-size cap of the interpolation error, the output-time grid (t0,
     * next frame, next D row), periodic shifts, the escape and end flags */
    int ipt, nf, nr, esc, ending;
    int jt;                     /* next trajectory output step */
    double told, hd, hopt_int, t0;
    double xs[3];
};

/* slot s of this lane's dense-output scratch */
#define DSV(a, s) ((a).ds[(size_t)(s) * (a).nl + gl])

__device__ __forceinline__ double bs_time(const BsArgs &a, const BsLane &L, long j)
{
    return L.t0 + (double)j * a.dt;
}

/* particle_bc, src/tracking.c:1172-1221, on the lane's state */
__device__ __forceinline__ void bs_particle_bc(const BsArgs &a, BsLane &L)
{
    if (a.bc == 0) {
        double d0 = a.hi[0] - a.lo[0], d1 = a.hi[1] - a.lo[1], d2 = a.hi[2] - a.lo[2];
        if (L.yv[0] > a.hi[0]) { L.yv[0] -= d0; L.xs[0] += d0; }
        if (L.yv[1] > a.hi[1]) { L.yv[1] -= d1; L.xs[1] += d1; }
        if (L.yv[2] > a.hi[2]) { L.yv[2] -= d2; L.xs[2] += d2; }
        if (L.yv[0] < a.lo[0]) { L.yv[0] += d0; L.xs[0] -= d0; }
        if (L.yv[1] < a.lo[1]) { L.yv[1] += d1; L.xs[1] -= d1; }
        if (L.yv[2] < a.lo[2]) { L.yv[2] += d2; L.xs[2] -= d2; }
    } else if (a.bc == 1) {
        if (L.yv[0] > a.hi[0] || L.yv[1] > a.hi[1] || L.yv[2] > a.hi[2] || L.yv[0] < a.lo[0] ||
            L.yv[1] < a.lo[1] || L.yv[2] < a.lo[2])
            L.esc = 1;
    }
}

/* ptl_energy_fixed's bin, src/diagnostics.c:187-205 */
__device__ __forceinline__ int bs_energy_bin(const BsArgs &a, double ux, double uy, double uz)
{
    double gama = sqrt(1.0 + ux * ux + uy * uy + uz * uz);
    double v = floor((log10(gama - 1.0) - a.logemin) / a.logde);
    if (!(v > -2147483649.0 && v < 2147483648.0)) return 0;
    int ib = (int)v;
    if (ib < 0) ib = 0;
    else if (ib > a.nbins - 1) ib = a.nbins - 1;
    return ib;
}

/* calc_diff_coeff's eight terms (src/diagnostics.c:659-705) of state y at
 * time tt with shifts xs and field e into row r, or, escaped, the stale
 * terms (src/tracking.c:334-336) */
__device__ __forceinline__ void bs_dterms(const BsArgs &a, long i, int r, int escaped,
                                          const double *y, double tt, const double *xs,
                                          const Emf &e)
{
    const long n = a.n;
    double *st = a.st;
    double dxx = st[i], dyy = st[n + i], dzz = st[2 * n + i];
    double daa = st[3 * n + i], dee = st[4 * n + i];
    double drr = 0.0, dpp = 0.0, duu = 0.0;
    if (escaped) {
        if (a.fix_stale) dxx = dyy = dzz = daa = dee = 0.0;
    } else {
        double t = tt * 2 + a.dt;
        double q;
        q = a.x0[i] - (y[0] + xs[0]); dxx = q * q / t;
        q = a.y0[i] - (y[1] + xs[1]); dyy = q * q / t;
        q = a.z0[i] - (y[2] + xs[2]); dzz = q * q / t;
        drr = dxx + dyy + dzz;
        double ax = a.ux0[i] - y[3], ay = a.uy0[i] - y[4], az = a.uz0[i] - y[5];
        dpp = (ax * ax + ay * ay + az * az) / t;
        double g = sqrt(1.0 + y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
        q = g - a.g0[i]; dee = q * q / t;
        double ab = sqrt(e.bx * e.bx + e.by * e.by + e.bz * e.bz);
        double au = sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
        double mu = (e.bx * y[3] + e.by * y[4] + e.bz * y[5]) / (ab * au);
        q = mu - a.mu0[i]; duu = q * q / t;
        double th = acos(mu) * 180 / TP_PI;
        q = th - a.th0[i]; daa = q * q / t;
    }
    st[i] = dxx; st[n + i] = dyy; st[2 * n + i] = dzz;
    st[3 * n + i] = daa; st[4 * n + i] = dee;
    double *o = a.dbuf + (size_t)r * 8 * n + i;
    o[0] = drr; o[n] = dxx; o[2 * n] = dyy; o[3 * n] = dzz;
    o[4 * n] = dpp; o[5 * n] = duu; o[6 * n] = daa; o[7 * n] = dee;
}

template <bool DENSE>
__device__ __forceinline__ void bs_finish(const BsArgs &a, BsLane &L, int failed, long gl)
{
    const long i = L.pi;
    a.x[i] = L.yv[0]; a.y[i] = L.yv[1]; a.z[i] = L.yv[2];
    a.ux[i] = L.yv[3]; a.uy[i] = L.yv[4]; a.uz[i] = L.yv[5];
    a.t[i] = L.t;
    a.nsteps[i] = L.acc;
    if (DENSE) {
        /* rows from nr on: the particle has left (or failed); it is counted
         * out from row 1 if it started outside, else from the row after its
         * first stale row (the fixed-step loop's count, src/tracking.c:366-369) */
        const Emf e = {};
        for (int r = L.nr; r < a.nd; r++) bs_dterms(a, i, r, 1, L.yv, 0.0, L.xs, e);
        if (L.esc) {
            const int from = L.ending < 0 ? 1 : L.nr + 1;
            if (from < a.nd) atomicAdd(&a.outhist[from], 1);
        }
        a.xs[i] = L.xs[0]; a.ys[i] = L.xs[1]; a.zs[i] = L.xs[2];
        a.flag[i] = L.esc ? 2 : failed;
    } else {
        a.flag[i] = failed;
    }
    atomicAdd(&a.ctr[Q_ACC], (unsigned long long)L.acc);
    atomicAdd(&a.ctr[Q_REJ], (unsigned long long)L.rej);
    atomicAdd(&a.ctr[Q_RHS], (unsigned long long)L.rhs);
    if (failed) atomicAdd(&a.ctr[Q_FAIL], 1ull);
    L.pi = -1;
}

__device__ __forceinline__ void bs_begin_seq(const BsArgs &a, BsLane &L)
{
    L.nstep = a.c.nseq[L.k];
    L.hsub = L.h / L.nstep;
    for (int i = 0; i < BS_N; i++) {
        L.ym[i] = L.yv[i];
        L.yn[i] = L.yv[i] + L.hsub * L.dydx[i];
    }
    L.xe = L.t + L.hsub;
    L.nn = 1;
    L.phase = PH_MID;
}

template <bool DENSE>
__device__ __forceinline__ void bs_begin_attempt(const BsArgs &a, BsLane &L, long gl)
{
    L.h = L.hnew;
    L.reject = 0;
    if (++L.attempts > a.cap) {
        bs_finish<DENSE>(a, L, 1, gl);
        return;
    }
    L.k = 0;
    L.ipt = 0;
    bs_begin_seq(a, L);
}

/* one trajectory point of the lane's particle: state y at time t with the
 * lane's shifts, folded into the box for periodic boundaries as particle_bc
 * folds the fixed-step loop's records (src/tracking.c:440-447, 1172-1221) */
__device__ __forceinline__ void bs_traj_record(const BsArgs &a, const BsLane &L, const double *y,
                                               double t)
{
    const long i = L.pi;
    const int toff = i > 0 ? a.traj_accum[i - 1] : 0;
    const int k = a.traj_pts[i];
    if (k < a.traj_accum[i] - toff) {
        tp_particle r;
        r.x = y[0]; r.y = y[1]; r.z = y[2]; r.qm = a.stage[i].qm;
        r.ux = y[3]; r.uy = y[4]; r.uz = y[5]; r.t = t;
        r.xshift = L.xs[0]; r.yshift = L.xs[1]; r.zshift = L.xs[2];
        if (a.bc == 0) {
            const double d0 = a.hi[0] - a.lo[0], d1 = a.hi[1] - a.lo[1], d2 = a.hi[2] - a.lo[2];
            if (r.x > a.hi[0]) { r.x -= d0; r.xshift += d0; }
            if (r.y > a.hi[1]) { r.y -= d1; r.yshift += d1; }
            if (r.z > a.hi[2]) { r.z -= d2; r.zshift += d2; }
            if (r.x < a.lo[0]) { r.x += d0; r.xshift -= d0; }
            if (r.y < a.lo[1]) { r.y += d1; r.yshift -= d1; }
            if (r.z < a.lo[2]) { r.z += d2; r.zshift -= d2; }
        }
        a.traj[toff + k] = r;
    }
    a.traj_pts[i] = k + 1;
}

/* the output times at the end time: the remaining frames from the final
 * state; the remaining D rows each need the field there (PH_DROW) */
template <bool DENSE>
__device__ __forceinline__ void bs_begin_end_outputs(const BsArgs &a, BsLane &L, long gl)
{
    for (; L.nf < a.nframes; L.nf++)
        atomicAdd(&a.espect[L.nf * a.nbins + bs_energy_bin(a, L.yv[3], L.yv[4], L.yv[5])], 1u);
    if (a.traj)
        for (; L.jt <= a.estep; L.jt += a.traj_nout) bs_traj_record(a, L, L.yv, bs_time(a, L, L.jt));
    if (L.nr < a.nd) {
        L.ending = 1;
        for (int i = 0; i < BS_N; i++) L.yn[i] = L.yv[i];
        L.xe = bs_time(a, L, (long)L.nr * a.di);
        L.phase = PH_DROW;
        return;
    }
    bs_finish<DENSE>(a, L, 0, gl);
}

/* top of step(): src/unused/StepperBS.c:173-191 */
template <bool DENSE>
__device__ __forceinline__ void bs_start_step(const BsArgs &a, BsLane &L, long gl)
{
    if (!(L.t < L.tend - 1e-13 * fabs(L.tend))) {
        if (DENSE) bs_begin_end_outputs<DENSE>(a, L, gl);
        else bs_finish<DENSE>(a, L, 0, gl);
        return;
    }
    double h = L.hnext;
    if (L.t + h > L.tend) h = L.tend - L.t;
    L.work[0] = 0.0;
    if (h != L.hnext && !L.first) L.last = 1;
    if (L.reject) {
        L.prev_reject = 1;
        L.last = 0;
    }
    L.reject = 0;
    L.hnew = fabs(h);
    bs_begin_attempt<DENSE>(a, L, gl);
}

/* after row k's midpoint sequence: extrapolation, error and the
 * convergence/rejection tests of src/unused/StepperBS.c:206-272 */
template <bool DENSE>
__device__ __forceinline__ void bs_end_seq(const BsArgs &a, BsLane &L, const double *yseq,
                                           long gl)
{
    const int k = L.k;
    if (k == 0) {
        for (int i = 0; i < BS_N; i++) L.ycur[i] = yseq[i];
        L.k = 1;
        bs_begin_seq(a, L);
        return;
    }
    for (int i = 0; i < BS_N; i++) L.tab[k - 1][i] = yseq[i];
    for (int j = k - 1; j > 0; j--)
        for (int i = 0; i < BS_N; i++)
            L.tab[j - 1][i] = L.tab[j][i] + a.c.coeff[k][j] * (L.tab[j][i] - L.tab[j - 1][i]);
    for (int i = 0; i < BS_N; i++)
        L.ycur[i] = L.tab[0][i] + a.c.coeff[k][0] * (L.tab[0][i] - L.ycur[i]);
    double err = 0.0;
    for (int i = 0; i < BS_N; i++) {
        double p = fabs(L.yv[i]), q = fabs(L.ycur[i]);
        double scale = a.atol + a.rtol * (p > q ? p : q);
        double r = (L.ycur[i] - L.tab[0][i]) / scale;
        err += r * r;
    }
    err = sqrt(err / BS_N);
    double expo = 1.0 / (2 * k + 1);
    double facmin = tp_pow(BS_STEPFAC3, expo);
    double fac;
    if (err == 0.0) {
        fac = 1.0 / facmin;
    } else {
        fac = BS_STEPFAC2 / tp_pow(err / BS_STEPFAC1, expo);
        double lo = facmin / BS_STEPFAC4, hi = 1.0 / facmin;
        fac = fac < hi ? fac : hi;
        fac = lo > fac ? lo : fac;
    }
    L.hopt[k] = fabs(L.h * fac);
    L.work[k] = a.c.cost[k] / L.hopt[k];
    bool brk = false;
    if ((L.first || L.last) && err <= 1.0) brk = true;
    if (!brk && k == L.ktarg - 1 && !L.prev_reject && !L.first && !L.last) {
        if (err <= 1.0) {
            brk = true;
        } else {
            double r = (double)(a.c.nseq[L.ktarg] * a.c.nseq[L.ktarg + 1] /
                                (a.c.nseq[0] * a.c.nseq[0]));
            if (err > r * r) {
                L.reject = 1;
                L.ktarg = k;
                if (L.ktarg > 1 && L.work[k - 1] < BS_KFAC1 * L.work[k]) L.ktarg--;
                L.hnew = L.hopt[L.ktarg];
                brk = true;
            }
        }
    }
    if (!brk && k == L.ktarg) {
        if (err <= 1.0) {
            brk = true;
        } else {
            double r = (double)a.c.nseq[k + 1] / a.c.nseq[0];
            if (err > r * r) {
                L.reject = 1;
                if (L.ktarg > 1 && L.work[k - 1] < BS_KFAC1 * L.work[k]) L.ktarg--;
                L.hnew = L.hopt[L.ktarg];
                brk = true;
            }
        }
    }
    if (!brk && k == L.ktarg + 1) {
        if (err > 1.0) {
            L.reject = 1;
            if (L.ktarg > 1 && L.work[L.ktarg - 1] < BS_KFAC1 * L.work[L.ktarg]) L.ktarg--;
            L.hnew = L.hopt[L.ktarg];
        }
        brk = true;
    }
    if (!brk) {
        L.k = k + 1;
        bs_begin_seq(a, L);
        return;
    }
    if (L.reject) {
        L.prev_reject = 1;
        L.rej++;
        bs_begin_attempt<DENSE>(a, L, gl);
        return;
    }
    /* accepted: the derivative at the new point comes next */
    L.xe = L.t + L.h;
    L.phase = PH_ACCEPT;
}

/* dense_interp, src/unused/StepperBS.c:478-524, on the lane's dens slots;
 * the coefficient array is indexed by compile-time orders (unrolled up to
 * 2 BS_KMAX - 1) so that it stays in registers */
__device__ __forceinline__ void bs_dense_interp(const BsArgs &a, long gl, int imit)
{
    const int n = BS_N;
    for (int i = 0; i < n; i++) {
        double y0 = DSV(a, DS_DENS + i), y1 = DSV(a, DS_DENS + 2 * n + i);
        double yp0 = DSV(a, DS_DENS + n + i), yp1 = DSV(a, DS_DENS + 3 * n + i);
        double ydiff = y1 - y0;
        double aspl = -yp1 + ydiff;
        double bspl = yp0 - ydiff;
        DSV(a, DS_DENS + n + i) = ydiff;
        DSV(a, DS_DENS + 2 * n + i) = aspl;
        DSV(a, DS_DENS + 3 * n + i) = bspl;
        if (imit < 0) continue;
        double ph0 = (y0 + y1) * 0.5 + 0.125 * (aspl + bspl);
        double ph1 = ydiff + (aspl - bspl) * 0.25;
        double ph2 = -(yp0 - yp1);
        double ph3 = 6.0 * (bspl - aspl);
        double av[2 * BS_KMAX];
        if (imit >= 1) {
            av[1] = 16.0 * (DSV(a, DS_DENS + 5 * n + i) - ph1);
            if (imit >= 3) {
                av[3] = 16.0 * (DSV(a, DS_DENS + 7 * n + i) - ph3 + 3 * av[1]);
#pragma unroll
                for (int im = 5; im <= 2 * BS_KMAX - 1; im += 2) {
                    if (im > imit) break;
                    double fac1 = im * (im - 1) / 2.0;
                    double fac2 = fac1 * (im - 2) * (im - 3) * 2.0;
                    av[im] = 16.0 * (DSV(a, DS_DENS + (im + 4) * n + i) + fac1 * av[im - 2] -
                                     fac2 * av[im - 4]);
                }
            }
        }
        av[0] = (DSV(a, DS_DENS + 4 * n + i) - ph0) * 16.0;
        if (imit >= 2) {
            av[2] = (DSV(a, DS_DENS + 6 * n + i) - ph2 + av[0]) * 16.0;
#pragma unroll
            for (int im = 4; im <= 2 * BS_KMAX - 1; im += 2) {
                if (im > imit) break;
                double fac1 = im * (im - 1) / 2.0;
                double fac2 = (double)(im * (im - 1) * (im - 2) * (im - 3));
                av[im] = 16.0 * (DSV(a, DS_DENS + (im + 4) * n + i) + fac1 * av[im - 2] -
                                 fac2 * av[im - 4]);
            }
        }
#pragma unroll
        for (int im = 0; im <= 2 * BS_KMAX - 1; im++) {
            if (im > imit) break;
            DSV(a, DS_DENS + n * (im + 4) + i) = av[im];
        }
    }
}

/* prepare_dense's polynomial extrapolation over sequences KB..k of one
 * component held in y: y[l-1] = y[l] + (y[l] - y[l-1]) / (r^2 - 1), r =
 * n_j / n_{l-1}, for j = KB+1..k and l = j..KB+1 (src/unused/StepperBS.c:
 * 545-553 and 564-572); returns y[KB] */
template <int KB>
__device__ __forceinline__ double bs_dense_extrap(const BsArgs &a, double (&y)[BS_KMAX + 1],
                                                  int k)
{
#pragma unroll
    for (int j = KB + 1; j <= BS_KMAX; j++) {
        if (j > k) break;
        double dblenj = a.c.nseq[j];
#pragma unroll
        for (int l = j; l >= KB + 1; l--) {
            double r = dblenj / a.c.nseq[l - 1];
            double factor = r * r - 1.0;
            y[l - 1] = y[l] + (y[l] - y[l - 1]) / factor;
        }
    }
    return y[KB];
}

/* sequence KK's derivative rows of component i (rows 2 KK^2 .. 2 (KK+1)^2 - 1
 * of fsave: ipoint[kk] = 2 kk^2 for nseq = 4k + 2) in registers, order by
 * order: the value order KMI takes from them (times (n_KK/2)^(KMI-1)) goes to
 * YV, then the differences prepare_dense takes between orders
 * (src/unused/StepperBS.c:558-601), element by element in its order.  Each
 * difference pairs rows two apart, so the rows at odd and at even offsets
 * in the sequence evolve separately: P selects one half (offset o is held
 * in row[o / 2]), which halves the registers held. */
template <int KK, int P, int KMI>
__device__ __forceinline__ void bs_dense_orders(const BsArgs &a, long gl, int mu, double dydx_i,
                                                double (&row)[2 * KK + 1])
{
    if constexpr (KMI <= 2 * BS_KMAX - 1) {
        constexpr int R = 4 * KK + 2;
        if (KMI > mu) return;
        if constexpr (KK >= (KMI - 1) / 2 && (2 * KK + KMI - 1) % 2 == P)
            DSV(a, DS_YV + (KMI - 1) * (BS_KMAX + 1) + KK) =
                row[(2 * KK + KMI - 1) / 2] * tp_powi(a.c.nseq[KK] / 2.0, KMI - 1);
        if (KMI == mu) return;
        if constexpr (KK >= KMI / 2) {
            if constexpr (P == 1) {
#pragma unroll
                for (int o = R - 1; o >= (KMI == 1 ? 3 : KMI); o -= 2)
                    row[o / 2] = row[o / 2] - row[(o - 2) / 2];
                if constexpr (KMI == 1) row[0] = row[0] - dydx_i;    /* offset 1 */
            } else {
#pragma unroll
                for (int o = R - 2; o >= KMI + 1; o -= 2)
                    row[o / 2] = row[o / 2] - row[(o - 2) / 2];
            }
        }
        bs_dense_orders<KK, P, KMI + 1>(a, gl, mu, dydx_i, row);
    }
}

template <int KK>
__device__ __forceinline__ void bs_dense_seq(const BsArgs &a, long gl, int i, int k, int mu,
                                             double dydx_i)
{
    if constexpr (KK <= BS_KMAX) {
        if (KK <= k) {
            constexpr int R0 = 2 * KK * KK;
#pragma unroll
            for (int P = 0; P < 2; P++) {
                double row[2 * KK + 1];
#pragma unroll
                for (int q = 0; q < 2 * KK + 1; q++)
                    row[q] = DSV(a, DS_FSAVE + (R0 + 2 * q + P) * BS_N + i);
                if (P == 0) bs_dense_orders<KK, 0, 1>(a, gl, mu, dydx_i, row);
                else bs_dense_orders<KK, 1, 1>(a, gl, mu, dydx_i, row);
            }
        }
        bs_dense_seq<KK + 1>(a, gl, i, k, mu, dydx_i);
    }
}

/* order KMI's dens coefficient of component i: the extrapolation of its
 * values over sequences (KMI-1)/2 .. k */
template <int KMI>
__device__ __forceinline__ void bs_dense_coeffs(const BsArgs &a, long gl, int i, int k, int mu,
                                                double h)
{
    if constexpr (KMI <= 2 * BS_KMAX - 1) {
        if (KMI > mu) return;
        constexpr int KB = (KMI - 1) / 2;
        double y[BS_KMAX + 1];
#pragma unroll
        for (int kk = KB; kk <= BS_KMAX; kk++) {
            if (kk > k) break;
            y[kk] = DSV(a, DS_YV + (KMI - 1) * (BS_KMAX + 1) + kk);
        }
        DSV(a, DS_DENS + (KMI + 4) * BS_N + i) = bs_dense_extrap<KB>(a, y, k) * h;
        bs_dense_coeffs<KMI + 1>(a, gl, i, k, mu, h);
    }
}

/* prepare_dense, src/unused/StepperBS.c:526-615: the dense-output values of
 * the accepted step of size h (start yv, end ycur, derivatives dydx and
 * dydxnew), extrapolated from the midpoint values and derivative differences
 * of sequences 0..k; returns the interpolation error estimate.  The
 * reference updates its ysave and fsave arrays in place order by order;
 * here each component's values go through the same operations in
 * registers, a sequence at a time, and the values each order needs pass
 * through the YV slots: the same results with a few hundred scratch
 * accesses per step instead of thousands of dependent ones. */
__device__ __forceinline__ double bs_prepare_dense(const BsArgs &a, const BsLane &L, long gl,
                                                   const double *dydxnew)
{
    const int n = BS_N, k = L.k, mu = 2 * k - 1;
    const double h = L.h;
    for (int i = 0; i < n; i++) {
        DSV(a, DS_DENS + i) = L.yv[i];
        DSV(a, DS_DENS + n + i) = h * L.dydx[i];
        DSV(a, DS_DENS + 2 * n + i) = L.ycur[i];
        DSV(a, DS_DENS + 3 * n + i) = dydxnew[i] * h;
    }
    for (int i = 0; i < n; i++) {
        /* the midpoint values, extrapolated */
        double y[BS_KMAX + 1];
#pragma unroll
        for (int kk = 0; kk <= BS_KMAX; kk++) {
            if (kk > k) break;
            y[kk] = DSV(a, DS_YSAVE + kk * BS_N + i);
        }
        DSV(a, DS_DENS + 4 * n + i) = bs_dense_extrap<0>(a, y, k);
        /* the derivative differences, order by order */
        bs_dense_seq<0>(a, gl, i, k, mu, L.dydx[i]);
        bs_dense_coeffs<1>(a, gl, i, k, mu, h);
    }
    bs_dense_interp(a, gl, mu);
    double error = 0.0;
    if (mu >= 1) {
        for (int i = 0; i < n; i++) {
            double p = fabs(L.yv[i]), q = fabs(L.ycur[i]);
            double scale = a.atol + a.rtol * (p > q ? p : q);
            double v = DSV(a, DS_DENS + (mu + 4) * n + i) / scale;
            error += v * v;
        }
        error = sqrt(error / n) * a.c.errfac[mu - 1];
    }
    return error;
}

/* dense_out, src/unused/StepperBS.c:617-644: component i at time x of the
 * lane's last accepted step */
__device__ __forceinline__ double bs_dense_out(const BsArgs &a, const BsLane &L, long gl, int i,
                                               double x)
{
    const int n = BS_N, mu = 2 * L.k - 1;
    double theta = (x - L.told) / L.hd;
    double theta1 = 1.0 - theta;
    double yinterp = DSV(a, DS_DENS + i) +
                     theta * (DSV(a, DS_DENS + n + i) +
                              theta1 * (DSV(a, DS_DENS + 2 * n + i) * theta +
                                        DSV(a, DS_DENS + 3 * n + i) * theta1));
    if (mu < 0) return yinterp;
    double theta05 = theta - 0.5;
    double tt = theta * theta1;
    double t4 = tt * tt;
    double cc = DSV(a, DS_DENS + n * (mu + 4) + i);
    for (int j = mu; j > 0; j--) cc = DSV(a, DS_DENS + n * (j + 3) + i) + cc * theta05 / j;
    return yinterp + t4 * cc;
}

/* the step's own output times (strictly before its end) next: a D row
 * needs the field at its point, evaluated in the next iteration */
template <bool DENSE>
__device__ __forceinline__ bool bs_next_drow(const BsArgs &a, BsLane &L, long gl)
{
    if (L.nr >= a.nd) return false;
    const double tr = bs_time(a, L, (long)L.nr * a.di);
    if (!(tr < L.told + L.hd)) return false;
    for (int i = 0; i < BS_N; i++) L.yn[i] = bs_dense_out(a, L, gl, i, tr);
    L.xe = tr;
    L.phase = PH_DROW;
    return true;
}

/* the accepted step's bookkeeping after its derivative f at the new point:
 * order and step-size choice of src/unused/StepperBS.c:296-347, then (dense
 * form) the boundary and the escape */
template <bool DENSE>
__device__ __forceinline__ void bs_accept_finish(const BsArgs &a, BsLane &L, const double *f,
                                                 long gl)
{
    for (int i = 0; i < BS_N; i++) {
        L.dydx[i] = f[i];
        L.yv[i] = L.ycur[i];
    }
    L.t += L.h;
    L.first = 0;
    L.acc++;
    const int k = L.k, KM1 = a.c.kmax - 1;
    int kopt;
    if (k == 1) {
        kopt = 2;
    } else if (k <= L.ktarg) {
        kopt = k;
        if (L.work[k - 1] < BS_KFAC1 * L.work[k]) kopt = k - 1;
        else if (L.work[k] < BS_KFAC2 * L.work[k - 1]) kopt = (k + 1 < KM1) ? k + 1 : KM1;
    } else {
        kopt = k - 1;
        if (k > 2 && L.work[k - 2] < BS_KFAC1 * L.work[k - 1]) kopt = k - 2;
        if (L.work[k] < BS_KFAC2 * L.work[kopt]) kopt = (k < KM1) ? k : KM1;
    }
    double hnew;
    if (L.prev_reject) {
        L.ktarg = kopt < k ? kopt : k;
        double p = fabs(L.h), q = L.hopt[L.ktarg];
        hnew = p < q ? p : q;
        L.prev_reject = 0;
    } else {
        if (kopt <= k) {
            hnew = L.hopt[kopt];
        } else {
            if (k < L.ktarg && L.work[k] < BS_KFAC2 * L.work[k - 1])
                hnew = L.hopt[k] * a.c.cost[kopt + 1] / a.c.cost[k];
            else
                hnew = L.hopt[k] * a.c.cost[kopt] / a.c.cost[k];
        }
        L.ktarg = kopt;
    }
    if (DENSE) {
        double hi = fabs(L.hopt_int);
        hnew = hnew < hi ? hnew : hi;
        L.hnext = hnew;
        bs_particle_bc(a, L);
        if (L.esc) {
            /* escape spectrum at the frame of the next output time */
            if (L.nf < a.nt_out)
                atomicAdd(&a.espect_esc[L.nf * a.nbins +
                                        bs_energy_bin(a, L.yv[3], L.yv[4], L.yv[5])], 1u);
            bs_finish<DENSE>(a, L, 0, gl);
            return;
        }
        bs_start_step<DENSE>(a, L, gl);
        return;
    }
    L.hnext = hnew;
    bs_start_step<DENSE>(a, L, gl);
}

/* after the derivative at the accepted point: the dense form first checks
 * the interpolation error (above 10 the step is redone with the step size
 * it suggests), bins the frames inside the step and queues its D rows */
template <bool DENSE>
__device__ __forceinline__ void bs_accept(const BsArgs &a, BsLane &L, const double *f, long gl)
{
    if (DENSE) {
        /* dydxnew waits in ym (free after the step) for the lane's turn */
        for (int i = 0; i < BS_N; i++) L.ym[i] = f[i];
        L.phase = PH_PREP;
        return;
    }
    bs_accept_finish<DENSE>(a, L, f, gl);
}

/* the dense form's rest of the accepted step, with dydxnew in ym */
template <bool DENSE>
__device__ __forceinline__ void bs_prep(const BsArgs &a, BsLane &L, long gl)
{
    const double *f = L.ym;
    {
        const double err = bs_prepare_dense(a, L, gl, f);
        const double pe = tp_pow(err, 1.0 / (2 * L.k + 3));
        L.hopt_int = L.h / (pe > 0.01 ? pe : 0.01);
        if (err > 10.0) {
            L.hnew = fabs(L.hopt_int);
            L.reject = 1;
            L.prev_reject = 1;
            L.rej++;
            bs_begin_attempt<DENSE>(a, L, gl);
            return;
        }
        L.told = L.t;
        L.hd = L.h;
        const double tnew = L.t + L.h;
        for (; L.nf < a.nframes; L.nf++) {
            const double tf = bs_time(a, L, (long)(L.nf + 1) * a.ei);
            if (!(tf < tnew)) break;
            const double ux = bs_dense_out(a, L, gl, 3, tf);
            const double uy = bs_dense_out(a, L, gl, 4, tf);
            const double uz = bs_dense_out(a, L, gl, 5, tf);
            atomicAdd(&a.espect[L.nf * a.nbins + bs_energy_bin(a, ux, uy, uz)], 1u);
        }
        /* trajectory points strictly inside the step, from the polynomial */
        if (a.traj) {
            for (; L.jt <= a.estep; L.jt += a.traj_nout) {
                const double tj = bs_time(a, L, L.jt);
                if (!(tj < tnew)) break;
                double yj[BS_N];
                for (int m = 0; m < BS_N; m++) yj[m] = bs_dense_out(a, L, gl, m, tj);
                bs_traj_record(a, L, yj, tj);
            }
        }
        /* ym keeps dydxnew while the D rows run */
        if (bs_next_drow<DENSE>(a, L, gl)) return;
    }
    bs_accept_finish<DENSE>(a, L, f, gl);
}

/* after the field e at a D row's point: its terms, then the next row or
 * the rest of the step (or, at the end time, of the particle) */
template <bool DENSE>
__device__ __forceinline__ void bs_drow(const BsArgs &a, BsLane &L, const Emf &e, long gl)
{
    bs_dterms(a, L.pi, L.nr, 0, L.yn, L.xe, L.xs, e);
    L.nr++;
    if (L.ending) {
        if (L.nr < a.nd) {
            L.xe = bs_time(a, L, (long)L.nr * a.di);
            return;                                    /* same point, next row */
        }
        bs_finish<DENSE>(a, L, 0, gl);
        return;
    }
    if (bs_next_drow<DENSE>(a, L, gl)) return;
    bs_accept_finish<DENSE>(a, L, L.ym, gl);
}

template <int FK, bool DENSE>
__global__ void __launch_bounds__(256) k_bs(BsArgs a)
{
    const int lane = threadIdx.x & 63;
    const long gl = (long)blockIdx.x * blockDim.x + threadIdx.x;
    BsLane L;
    L.pi = -1;
    while (true) {
        /* refill idle lanes: one atomic per wavefront */
        unsigned long long need = __ballot(L.pi == -1);
        if (need) {
            int leader = __ffsll((long long)need) - 1;
            unsigned cnt = (unsigned)__popcll(need);
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(&a.ctr[Q_HEAD], (unsigned long long)cnt);
            unsigned lo = __shfl((unsigned)(base & 0xffffffffull), leader, 64);
            unsigned hi = __shfl((unsigned)(base >> 32), leader, 64);
            base = ((unsigned long long)hi << 32) | lo;
            if (L.pi == -1) {
                unsigned long long below = lane ? (need & ((1ull << lane) - 1ull)) : 0ull;
                unsigned long long idx = base + (unsigned long long)__popcll(below);
                if (idx < (unsigned long long)a.n) {
                    const long i = (long)idx;
                    L.pi = i;
                    L.yv[0] = a.x[i]; L.yv[1] = a.y[i]; L.yv[2] = a.z[i];
                    L.yv[3] = a.ux[i]; L.yv[4] = a.uy[i]; L.yv[5] = a.uz[i];
                    L.t = a.t[i];
                    double g0 = sqrt(1.0 + L.yv[3] * L.yv[3] + L.yv[4] * L.yv[4] +
                                     L.yv[5] * L.yv[5]);
                    double T = 2.0 * TP_PI * g0 / (fabs(a.qm) * a.b_ref);
                    L.tend = a.t_end_gyro > 0.0 ? L.t + a.t_end_gyro * T : a.t_end;
                    L.hnext = a.h0 > 0.0 ? a.h0 : 0.01 * T;
                    L.ktarg = a.c.ktarg0;
                    L.first = 1;
                    L.last = L.reject = L.prev_reject = 0;
                    L.attempts = L.acc = L.rej = 0;
                    L.rhs = 0;
                    L.xe = L.t;
                    L.phase = PH_INIT;
                    if (DENSE) {
                        L.t0 = L.t;
                        L.tend = L.t + a.estep * a.dt;
                        L.nf = 0; L.nr = 1; L.esc = 0; L.ending = 0;
                        L.jt = a.traj_nout;
                        L.xs[0] = a.xs[i]; L.xs[1] = a.ys[i]; L.xs[2] = a.zs[i];
                        /* initial wrap and escape check, src/tracking.c:314-332 */
                        bs_particle_bc(a, L);
                        if (L.esc) {
                            L.ending = -1;
                            bs_finish<DENSE>(a, L, 0, gl);
                        }
                    }
                } else {
                    L.pi = -2;
                }
            }
        }
        if (__ballot(L.pi >= 0) == 0ull) break;
        /* the one field evaluation of this iteration, executed by the whole
         * wavefront (retired lanes evaluate a dummy point) so the field sum
         * stays in uniform control flow and its uniform table reads become
         * scalar loads */
        const bool act = L.pi >= 0 && !(DENSE && L.phase == PH_PREP);
        double ze[BS_N];
        for (int i = 0; i < BS_N; i++)
            ze[i] = !act ? 0.0
                         : (L.phase == PH_INIT ? L.yv[i]
                                               : (L.phase == PH_ACCEPT ? L.ycur[i] : L.yn[i]));
        Emf e;
        get_emf64<FK>(a.f, ze[0], ze[1], ze[2], act ? L.xe : 0.0, e);
        if (act && DENSE && L.phase == PH_DROW) {
            bs_drow<DENSE>(a, L, e, gl);
        } else if (act) {
            double f[BS_N];
            bs_deriv(e, a.qm, ze, f);
            L.rhs++;
            if (L.phase == PH_INIT) {
                for (int i = 0; i < BS_N; i++) L.dydx[i] = f[i];
                bs_start_step<DENSE>(a, L, gl);
            } else if (L.phase == PH_MID) {
                if (L.nn < L.nstep) {
                    if (DENSE) {
                        if (L.nn == L.nstep / 2)
                            for (int i = 0; i < BS_N; i++)
                                DSV(a, DS_YSAVE + L.k * BS_N + i) = L.yn[i];
                        if (abs(L.nn - L.nstep / 2) <= 2 * L.k + 1) {
                            for (int i = 0; i < BS_N; i++)
                                DSV(a, DS_FSAVE + L.ipt * BS_N + i) = f[i];
                            L.ipt++;
                        }
                    }
                    double h2 = 2.0 * L.hsub;
                    for (int i = 0; i < BS_N; i++) {
                        double swap = L.ym[i] + h2 * f[i];
                        L.ym[i] = L.yn[i];
                        L.yn[i] = swap;
                    }
                    L.xe += L.hsub;
                    L.nn++;
                } else {
                    if (DENSE && L.nstep / 2 <= 2 * L.k + 1) {
                        for (int i = 0; i < BS_N; i++)
                            DSV(a, DS_FSAVE + L.ipt * BS_N + i) = f[i];
                        L.ipt++;
                    }
                    double yseq[BS_N];
                    for (int i = 0; i < BS_N; i++)
                        yseq[i] = 0.5 * (L.ym[i] + L.yn[i] + L.hsub * f[i]);
                    bs_end_seq<DENSE>(a, L, yseq, gl);
                }
            } else {
                bs_accept<DENSE>(a, L, f, gl);
            }
        }
        if (DENSE) {
            const unsigned long long pend = __ballot(L.pi >= 0 && L.phase == PH_PREP);
            if (pend) {
                const unsigned long long busy = __ballot(L.pi >= 0 && L.phase != PH_PREP);
                if ((__popcll(pend) >= BS_PREP_BATCH || busy == 0ull) && L.pi >= 0 &&
                    L.phase == PH_PREP)
                    bs_prep<DENSE>(a, L, gl);
            }
        }
    }
}

/* counter reset as a kernel (graph-safe, no memset node) */
__global__ void k_bs_reset(unsigned long long *ctr)
{
    if (threadIdx.x < Q_N) ctr[threadIdx.x] = 0ull;
}

typedef void (*bs_fn)(dim3, dim3, hipStream_t, const BsArgs &);

template <int FK, bool DENSE>
static void launch_bs(dim3 g, dim3 b, hipStream_t s, const BsArgs &a)
{
    hipLaunchKernelGGL((k_bs<FK, DENSE>), g, b, 0, s, a);
}

template <bool DENSE>
static bs_fn pick_bs_t(int fk)
{
    switch (fk) {
    case FK_TEST: return launch_bs<FK_TEST, DENSE>;
    case FK_ABC: return launch_bs<FK_ABC, DENSE>;
    case FK_ABCV: return launch_bs<FK_ABCV, DENSE>;
    case FK_GRID_REF: return launch_bs<FK_GRID_REF, DENSE>;
    case FK_GRID_TB: return launch_bs<FK_GRID_TB, DENSE>;
    case FK_MODES: return launch_bs<FK_MODES, DENSE>;
    case FK_WLCS: return launch_bs<FK_WLCS, DENSE>;
    case FK_MHD_REF: return launch_bs<FK_MHD_REF, DENSE>;
    case FK_GRIDD_REF: return launch_bs<FK_GRIDD_REF, DENSE>;
    }
    return nullptr;
}

static bs_fn pick_bs(int fk) { return pick_bs_t<false>(fk); }

/* D sums of the dense form: column (component q, row r) of the per-particle
 * terms summed over the particles in a fixed order (strided partial sums,
 * then an LDS tree), one block per column; row 0 stays 0 */
__global__ void __launch_bounds__(256) k_bs_dsum(const double *dbuf, long n, int nd,
                                                 double *dco)
{
    __shared__ double sh[256];
    const int q = blockIdx.x % 8, r = blockIdx.x / 8;
    const double *col = dbuf + ((size_t)r * 8 + q) * n;
    double s = 0.0;
    if (r > 0)
        for (long i = threadIdx.x; i < n; i += 256) s += col[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) dco[(size_t)q * nd + r] = sh[0];
}

/* escapes counted against each row: the particles counted out from row k
 * or before (outhist[k] counts those counted out from row k) */
__global__ void k_bs_rdec(const int *outhist, int nd, int *rdec)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int run = 0;
    for (int k = 0; k < nd; k++) {
        run += outhist[k];
        rdec[k] = run;
    }
}

__global__ void __launch_bounds__(256) k_bs_zero_u32(unsigned *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        p[i] = 0u;
}

void tp_bs_release(tp_ctx *c)
{
    if (c->bs_graph) {
        (void)hipGraphExecDestroy(c->bs_graph);
        c->bs_graph = nullptr;
    }
    (void)hipFree(c->bs_ctr);
    c->bs_ctr = nullptr;
    c->bs_graph_key = 0;
    (void)hipFree(c->bs_ds);
    (void)hipFree(c->bs_dbuf);
    (void)hipFree(c->bs_outhist);
    c->bs_ds = c->bs_dbuf = nullptr;
    c->bs_outhist = nullptr;
    c->bs_ds_cap = c->bs_dbuf_cap = c->bs_outhist_cap = 0;
}

static unsigned long long fnv(unsigned long long h, const void *p, size_t n)
{
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

static int bs_check(const tp_ctx *c, const tp_bs_params *p)
{
    if (!p || !(p->rtol > 0.0) || !(p->atol >= 0.0) || p->kmax < 3 || p->kmax > 8 ||
        !(p->qm != 0.0) || p->max_attempts < 0)
        return TP_E_ARG;
    if (p->t_end_gyro > 0.0 && !(p->b_ref > 0.0)) return TP_E_ARG;
    if (p->block != 0 && p->block != 64 && p->block != 128 && p->block != 256) return TP_E_ARG;
    if (p->waves_per_cu < 0 || p->waves_per_cu > 32) return TP_E_ARG;
    if (!c->have_field) return TP_E_STATE;
    return TP_OK;
}

static void bs_fill_args(const tp_ctx *c, const tp_bs_params *p, BsArgs &a, int n, double *x,
                         double *y, double *z, double *ux, double *uy, double *uz, double *t,
                         int *nsteps, int *flag, unsigned long long *ctr, int cap, int dense = 0)
{
    memset(&a, 0, sizeof(a));
    a.f = c->f;
    bs_consts(p->kmax, p->rtol, dense, &a.c);
    a.n = n;
    a.cap = cap;
    a.qm = p->qm; a.rtol = p->rtol; a.atol = p->atol;
    a.t_end = p->t_end; a.t_end_gyro = p->t_end_gyro; a.b_ref = p->b_ref; a.h0 = p->h0;
    a.x = x; a.y = y; a.z = z; a.ux = ux; a.uy = uy; a.uz = uz; a.t = t;
    a.nsteps = nsteps; a.flag = flag; a.ctr = ctr;
}

static int bs_ensure_ctr(tp_ctx *c)
{
    if (!c->bs_ctr && hipMalloc(&c->bs_ctr, sizeof(unsigned long long) * Q_N) != hipSuccess) {
        c->bs_ctr = nullptr;
        return TP_E_NOMEM;
    }
    return TP_OK;
}

static dim3 bs_grid(const tp_ctx *c, int block, int wpc, int n)
{
    int cu = c->num_cu > 0 ? c->num_cu : 256;
    long lanes = (long)cu * wpc * 64;
    long need = ((long)n + 63) / 64 * 64;       /* no more lanes than particles */
    if (lanes > need) lanes = need;
    long blocks = (lanes + block - 1) / block;
    if (blocks < 1) blocks = 1;
    return dim3((unsigned)blocks);
}

/* grow a device buffer of a context to at least `need` elements */
template <typename T>
static int bs_grow(T **buf, size_t *cap, size_t need)
{
    if (need <= *cap) return TP_OK;
    (void)hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    if (hipMalloc(buf, need * sizeof(T)) != hipSuccess) {
        *buf = nullptr;
        return TP_E_NOMEM;
    }
    *cap = need;
    return TP_OK;
}

extern "C" {

int tp_gpu_run_adaptive(tp_ctx *c, const tp_bs_params *p, tp_bs_stats *stats)
{
    if (!c) return TP_E_ARG;
    int rc = bs_check(c, p);
    if (rc) return rc;
    if (!c->s.derived && c->s.n > 0) return TP_E_STATE;
    bs_fn fn = pick_bs(c->f.kind);
    if (!fn) return TP_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    if ((rc = bs_ensure_ctr(c))) return rc;
    const int n = c->s.n;
    int block = p->block, wpc = p->waves_per_cu;
    if (block == 0 || wpc == 0) {
        int tb = 64, tw = 8;
        if (c->bs_tune_block) { tb = c->bs_tune_block[0]; tw = c->bs_tune_block[1]; }
        if (block == 0) block = tb;
        if (wpc == 0) wpc = tw;
    }
    const int cap = p->max_attempts > 0 ? p->max_attempts : 10000000;
    BsArgs a;
    State &S = c->s;
    bs_fill_args(c, p, a, n, S.x, S.y, S.z, S.ux, S.uy, S.uz, S.t, S.nsteps, S.esc, c->bs_ctr,
                 cap);
    dim3 grid = bs_grid(c, block, wpc, n);
    if (n > 0) {
        if (p->use_graph) {
            unsigned long long key = 1469598103934665603ull;
            key = fnv(key, &a, sizeof(a));
            key = fnv(key, &grid, sizeof(grid));
            key = fnv(key, &block, sizeof(block));
            if (!c->bs_graph || c->bs_graph_key != key) {
                if (c->bs_graph) {
                    (void)hipGraphExecDestroy(c->bs_graph);
                    c->bs_graph = nullptr;
                }
                hipGraph_t g;
                HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                hipLaunchKernelGGL(k_bs_reset, dim3(1), dim3(64), 0, c->stream, c->bs_ctr);
                fn(grid, dim3(block), c->stream, a);
                HIPCHK(hipStreamEndCapture(c->stream, &g));
                hipError_t e = hipGraphInstantiate(&c->bs_graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e != hipSuccess) {
                    c->bs_graph = nullptr;
                    return TP_E_HIP;
                }
                c->bs_graph_key = key;
            }
            HIPCHK(hipEventRecord(c->ev0, c->stream));
            HIPCHK(hipGraphLaunch(c->bs_graph, c->stream));
            HIPCHK(hipEventRecord(c->ev1, c->stream));
        } else {
            hipLaunchKernelGGL(k_bs_reset, dim3(1), dim3(64), 0, c->stream, c->bs_ctr);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(c->ev0, c->stream));
            fn(grid, dim3(block), c->stream, a);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(c->ev1, c->stream));
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        if (n > 0) {
            unsigned long long h[Q_N];
            HIPCHK(hipMemcpy(h, c->bs_ctr, sizeof(h), hipMemcpyDeviceToHost));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            stats->accepted = (long long)h[Q_ACC];
            stats->rejected = (long long)h[Q_REJ];
            stats->rhs_evals = (long long)h[Q_RHS];
            stats->failed = (int)h[Q_FAIL];
            stats->launches = 1;
            stats->kernel_ms = ms;
        }
    }
    return TP_OK;
}

int tp_gpu_run_adaptive_diag(tp_ctx *c, const tp_bs_params *p, const tp_track_params *q,
                             tp_bs_stats *stats)
{
    if (!c || !q) return TP_E_ARG;
    int rc = bs_check(c, p);
    if (rc) return rc;
    if (q->sstep != 1 || q->estep < 1 || q->einterval_t <= 0 || q->dcoeff_interval <= 0 ||
        q->nbins < 1 || q->nt_out < 1 || !(q->dt > 0.0) || q->bc_flag < 0 || q->bc_flag > 2 ||
        q->precision != TP_FP64)
        return TP_E_ARG;
    if ((size_t)q->nbins * sizeof(unsigned) > 64 * 1024) return TP_E_ARG;
    if (!c->s.derived) return TP_E_STATE;
    bs_fn fn = pick_bs_t<true>(c->f.kind);
    if (!fn) return TP_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    if ((rc = bs_ensure_ctr(c))) return rc;
    const int n = c->s.n;
    const int nd = q->estep / q->dcoeff_interval + 1;
    int nframes = q->estep / q->einterval_t;
    if (nframes > q->nt_out) nframes = q->nt_out;
    int block = p->block, wpc = p->waves_per_cu;
    if (block == 0 || wpc == 0) {
        int tb = 64, tw = 8;
        if (c->bs_tune_block) { tb = c->bs_tune_block[0]; tw = c->bs_tune_block[1]; }
        if (block == 0) block = tb;
        if (wpc == 0) wpc = tw;
    }
    const dim3 grid = bs_grid(c, block, wpc, n);
    const long nl = (long)grid.x * block;
    /* the run's outputs in the context's diagnostics buffers (read back by
     * tp_gpu_read_diagnostics), the scratch of the dense form */
    const size_t need = (size_t)q->nbins * q->nt_out;
    if (need > c->diag_cap) {
        (void)hipFree(c->espect); (void)hipFree(c->espect_esc);
        c->espect = nullptr; c->espect_esc = nullptr; c->diag_cap = 0;
        if (hipMalloc(&c->espect, need * sizeof(unsigned)) != hipSuccess ||
            hipMalloc(&c->espect_esc, need * sizeof(unsigned)) != hipSuccess)
            return TP_E_NOMEM;
        c->diag_cap = need;
    }
    if ((size_t)nd > c->dco_cap) {
        (void)hipFree(c->dco); (void)hipFree(c->rdec);
        c->dco = nullptr; c->rdec = nullptr; c->dco_cap = 0;
        if (hipMalloc(&c->dco, sizeof(double) * 8 * nd) != hipSuccess ||
            hipMalloc(&c->rdec, sizeof(int) * (nd + 1)) != hipSuccess)
            return TP_E_NOMEM;
        c->dco_cap = nd;
    }
    const size_t dneed = (size_t)nd * 8 * (n > 0 ? n : 1);
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess &&
        (dneed + (size_t)nl * DS_SLOTS) * sizeof(double) > fr / 2 + c->bs_dbuf_cap * sizeof(double))
        return TP_E_NOMEM;
    if ((rc = bs_grow(&c->bs_dbuf, &c->bs_dbuf_cap, dneed)) ||
        (rc = bs_grow(&c->bs_ds, &c->bs_ds_cap, (size_t)nl * DS_SLOTS)) ||
        (rc = bs_grow(&c->bs_outhist, &c->bs_outhist_cap, (size_t)nd)))
        return rc;
    hipLaunchKernelGGL(k_bs_zero_u32, dim3(256), dim3(256), 0, c->stream, c->espect, need);
    hipLaunchKernelGGL(k_bs_zero_u32, dim3(256), dim3(256), 0, c->stream, c->espect_esc, need);
    hipLaunchKernelGGL(k_bs_zero_u32, dim3(64), dim3(256), 0, c->stream,
                       (unsigned *)c->bs_outhist, (size_t)nd);
    HIPCHK(hipGetLastError());
    c->nb = q->nbins; c->nt = q->nt_out; c->nd = nd; c->last_n = n;
    const int cap = p->max_attempts > 0 ? p->max_attempts : 10000000;
    BsArgs a;
    State &S = c->s;
    bs_fill_args(c, p, a, n, S.x, S.y, S.z, S.ux, S.uy, S.uz, S.t, S.nsteps, S.esc, c->bs_ctr,
                 cap, 1);
    a.ds = c->bs_ds; a.nl = nl;
    a.dt = q->dt; a.estep = q->estep; a.ei = q->einterval_t; a.di = q->dcoeff_interval;
    a.nd = nd; a.nframes = nframes; a.nbins = q->nbins; a.nt_out = q->nt_out;
    a.bc = q->bc_flag; a.fix_stale = (q->flags & TP_FLAG_FIX_STALE_DCOEFF) ? 1 : 0;
    a.lo[0] = q->xmin; a.lo[1] = q->ymin; a.lo[2] = q->zmin;
    a.hi[0] = q->xmax; a.hi[1] = q->ymax; a.hi[2] = q->zmax;
    a.logemin = q->logemin; a.logde = q->logde;
    a.xs = S.xs; a.ys = S.ys; a.zs = S.zs;
    a.x0 = S.x0; a.y0 = S.y0; a.z0 = S.z0; a.ux0 = S.ux0; a.uy0 = S.uy0; a.uz0 = S.uz0;
    a.g0 = S.g0; a.mu0 = S.mu0; a.th0 = S.th0; a.st = S.st;
    a.espect = c->espect; a.espect_esc = c->espect_esc;
    a.dbuf = c->bs_dbuf; a.outhist = c->bs_outhist;
    if (c->traj_nout > 0) {
        /* armed by tp_gpu_track_adaptive_diag_traj */
        a.traj = c->traj; a.traj_accum = c->traj_accum; a.traj_pts = c->traj_pts;
        a.traj_nout = c->traj_nout; a.stage = S.stage;
    }
    hipLaunchKernelGGL(k_bs_reset, dim3(1), dim3(64), 0, c->stream, c->bs_ctr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    if (n > 0) fn(grid, dim3(block), c->stream, a);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    hipLaunchKernelGGL(k_bs_dsum, dim3(8 * nd), dim3(256), 0, c->stream,
                       (const double *)c->bs_dbuf, (long)n, nd, c->dco);
    hipLaunchKernelGGL(k_bs_rdec, dim3(1), dim3(64), 0, c->stream,
                       (const int *)c->bs_outhist, nd, c->rdec);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    unsigned long long h[Q_N];
    HIPCHK(hipMemcpy(h, c->bs_ctr, sizeof(h), hipMemcpyDeviceToHost));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = 1;
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->accepted = (long long)h[Q_ACC];
        stats->rejected = (long long)h[Q_REJ];
        stats->rhs_evals = (long long)h[Q_RHS];
        stats->failed = (int)h[Q_FAIL];
        stats->launches = 1;
        stats->kernel_ms = ms;
    }
    return TP_OK;
}

int tp_gpu_track_adaptive(tp_ctx *c, const tp_bs_params *p, int n, tp_particle *ptl,
                          int *nsteps, tp_bs_stats *stats)
{
    if (!c || !p || n < 0 || (n > 0 && !ptl)) return TP_E_ARG;
    int rc = bs_check(c, p);
    if (rc) return rc;
    if ((rc = tp_gpu_load_particles(c, n, ptl, ptl))) return rc;
    if ((rc = tp_gpu_run_adaptive(c, p, stats))) return rc;
    return tp_gpu_store_particles(c, ptl, nsteps);
}

int tp_gpu_tune_adaptive(tp_ctx *c, const tp_bs_params *p, const char *cache_path, int *block,
                         int *waves_per_cu)
{
    if (!c) return TP_E_ARG;
    int rc = bs_check(c, p);
    if (rc) return rc;
    if (!c->s.derived || c->s.n <= 0) return TP_E_STATE;
    bs_fn fn = pick_bs(c->f.kind);
    if (!fn) return TP_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    char key[160];
    snprintf(key, sizeof(key), "fk=%d modes=%d rtol=%.3g kmax=%d cu=%d", c->f.kind,
             c->f.nmodes, p->rtol, p->kmax, c->num_cu);
    int best_b = 0, best_w = 0;
    if (cache_path) {
        FILE *fp = fopen(cache_path, "r");
        if (fp) {
            char line[256];
            while (fgets(line, sizeof(line), fp)) {
                int b, w;
                double r;
                char k2[160];
                if (sscanf(line, "%d %d %lf %159[^\n]", &b, &w, &r, k2) == 4 &&
                    strcmp(k2, key) == 0) {
                    best_b = b;
                    best_w = w;
                }
            }
            fclose(fp);
        }
    }
    double best_rate = -1.0;
    if (!best_b) {
        if ((rc = bs_ensure_ctr(c))) return rc;
        const int nt = c->s.n < 65536 ? c->s.n : 65536;
        /* scratch copy of the first nt particles: 7 state arrays + 2 int */
        double *buf = nullptr;
        int *ib = nullptr;
        if (hipMalloc(&buf, sizeof(double) * 7 * nt) != hipSuccess) return TP_E_NOMEM;
        if (hipMalloc(&ib, sizeof(int) * 2 * nt) != hipSuccess) {
            (void)hipFree(buf);
            return TP_E_NOMEM;
        }
        const double *src[7] = {c->s.x, c->s.y, c->s.z, c->s.ux, c->s.uy, c->s.uz, c->s.t};
        const int blocks[3] = {64, 128, 256};
        const int wpcs[4] = {4, 8, 16, 24};
        for (int bi = 0; bi < 3 && rc == TP_OK; bi++) {
            for (int wi = 0; wi < 4 && rc == TP_OK; wi++) {
                for (int k = 0; k < 7; k++) {
                    if (hipMemcpyAsync(buf + (size_t)k * nt, src[k], sizeof(double) * nt,
                                       hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
                        rc = TP_E_HIP;
                }
                BsArgs a;
                /* bounded trial: 24 step attempts per particle */
                bs_fill_args(c, p, a, nt, buf, buf + nt, buf + 2 * (size_t)nt,
                             buf + 3 * (size_t)nt, buf + 4 * (size_t)nt, buf + 5 * (size_t)nt,
                             buf + 6 * (size_t)nt, ib, ib + nt, c->bs_ctr, 24);
                hipLaunchKernelGGL(k_bs_reset, dim3(1), dim3(64), 0, c->stream, c->bs_ctr);
                (void)hipEventRecord(c->ev0, c->stream);
                fn(bs_grid(c, blocks[bi], wpcs[wi], nt), dim3(blocks[bi]), c->stream, a);
                (void)hipEventRecord(c->ev1, c->stream);
                if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
                    rc = TP_E_HIP;
                unsigned long long h[Q_N];
                float ms = 0.f;
                if (rc == TP_OK &&
                    (hipMemcpy(h, c->bs_ctr, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess ||
                     hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess))
                    rc = TP_E_HIP;
                if (rc == TP_OK && ms > 0.f) {
                    double rate = (double)h[Q_RHS] / ms;
                    if (rate > best_rate) {
                        best_rate = rate;
                        best_b = blocks[bi];
                        best_w = wpcs[wi];
                    }
                }
            }
        }
        (void)hipFree(buf);
        (void)hipFree(ib);
        if (rc) return rc;
        if (cache_path && best_b) {
            FILE *fp = fopen(cache_path, "a");
            if (fp) {
                fprintf(fp, "%d %d %.6g %s\n", best_b, best_w, best_rate, key);
                fclose(fp);
            }
        }
    }
    if (!best_b) return TP_E_HIP;
    if (!c->bs_tune_block) {
        c->bs_tune_block = (int *)malloc(2 * sizeof(int));
        if (!c->bs_tune_block) return TP_E_NOMEM;
    }
    c->bs_tune_block[0] = best_b;
    c->bs_tune_block[1] = best_w;
    if (block) *block = best_b;
    if (waves_per_cu) *waves_per_cu = best_w;
    return TP_OK;
}

} /* extern "C" */
