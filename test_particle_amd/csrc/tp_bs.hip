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

enum { PH_INIT = 0, PH_MID = 1, PH_ACCEPT = 2 };
enum { Q_HEAD = 0, Q_ACC = 1, Q_REJ = 2, Q_RHS = 3, Q_FAIL = 4, Q_N = 8 };

struct BsConsts {
    int kmax, ktarg0;
    int nseq[9], cost[9];
    double coeff[9][9];
};

struct BsArgs {
    FieldDev f;
    BsConsts c;
    int n, cap;
    double qm, rtol, atol, t_end, t_end_gyro, b_ref, h0;
    double *x, *y, *z, *ux, *uy, *uz, *t;
    int *nsteps, *flag;
    unsigned long long *ctr;
};

/* StepperBase (src/unused/StepperBS.c:103-131), non-dense sequence */
static void bs_consts(int kmax, double rtol, BsConsts *c)
{
    memset(c, 0, sizeof(*c));
    c->kmax = kmax;
    for (int i = 0; i <= kmax; i++) c->nseq[i] = 2 * (i + 1);
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
}

template <int FK>
__device__ __forceinline__ void bs_rhs(const FieldDev &f, double qm, double t, const double *y,
                                       double *dy)
{
    Emf e;
    get_emf64<FK>(f, y[0], y[1], y[2], t, e);
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
};

__device__ __forceinline__ void bs_finish(const BsArgs &a, BsLane &L, int failed)
{
    const long i = L.pi;
    a.x[i] = L.yv[0]; a.y[i] = L.yv[1]; a.z[i] = L.yv[2];
    a.ux[i] = L.yv[3]; a.uy[i] = L.yv[4]; a.uz[i] = L.yv[5];
    a.t[i] = L.t;
    a.nsteps[i] = L.acc;
    a.flag[i] = failed;
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

__device__ __forceinline__ void bs_begin_attempt(const BsArgs &a, BsLane &L)
{
    L.h = L.hnew;
    L.reject = 0;
    if (++L.attempts > a.cap) {
        bs_finish(a, L, 1);
        return;
    }
    L.k = 0;
    bs_begin_seq(a, L);
}

/* top of step(): src/unused/StepperBS.c:173-191 */
__device__ __forceinline__ void bs_start_step(const BsArgs &a, BsLane &L)
{
    if (!(L.t < L.tend - 1e-13 * fabs(L.tend))) {
        bs_finish(a, L, 0);
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
    bs_begin_attempt(a, L);
}

/* after row k's midpoint sequence: extrapolation, error and the
 * convergence/rejection tests of src/unused/StepperBS.c:206-272 */
__device__ __forceinline__ void bs_end_seq(const BsArgs &a, BsLane &L, const double *yseq)
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
        bs_begin_attempt(a, L);
        return;
    }
    /* accepted: the derivative at the new point comes next */
    L.xe = L.t + L.h;
    L.phase = PH_ACCEPT;
}

/* after the derivative at the accepted point: order and step-size choice of
 * src/unused/StepperBS.c:296-347 */
__device__ __forceinline__ void bs_accept(const BsArgs &a, BsLane &L, const double *f)
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
    L.hnext = hnew;
    bs_start_step(a, L);
}

template <int FK>
__global__ void __launch_bounds__(256) k_bs(BsArgs a)
{
    const int lane = threadIdx.x & 63;
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
                } else {
                    L.pi = -2;
                }
            }
        }
        if (__ballot(L.pi >= 0) == 0ull) break;
        /* the one right-hand-side evaluation of this iteration, executed by
         * the whole wavefront (retired lanes evaluate a dummy point) so the
         * field sum stays in uniform control flow and its uniform table
         * reads become scalar loads */
        const bool act = L.pi >= 0;
        double ze[BS_N];
        for (int i = 0; i < BS_N; i++)
            ze[i] = !act ? 0.0 : (L.phase == PH_INIT ? L.yv[i] : (L.phase == PH_MID ? L.yn[i] : L.ycur[i]));
        double f[BS_N];
        bs_rhs<FK>(a.f, a.qm, act ? L.xe : 0.0, ze, f);
        if (act) {
            L.rhs++;
            if (L.phase == PH_INIT) {
                for (int i = 0; i < BS_N; i++) L.dydx[i] = f[i];
                bs_start_step(a, L);
            } else if (L.phase == PH_MID) {
                if (L.nn < L.nstep) {
                    double h2 = 2.0 * L.hsub;
                    for (int i = 0; i < BS_N; i++) {
                        double swap = L.ym[i] + h2 * f[i];
                        L.ym[i] = L.yn[i];
                        L.yn[i] = swap;
                    }
                    L.xe += L.hsub;
                    L.nn++;
                } else {
                    double yseq[BS_N];
                    for (int i = 0; i < BS_N; i++)
                        yseq[i] = 0.5 * (L.ym[i] + L.yn[i] + L.hsub * f[i]);
                    bs_end_seq(a, L, yseq);
                }
            } else {
                bs_accept(a, L, f);
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

template <int FK>
static void launch_bs(dim3 g, dim3 b, hipStream_t s, const BsArgs &a)
{
    hipLaunchKernelGGL((k_bs<FK>), g, b, 0, s, a);
}

static bs_fn pick_bs(int fk)
{
    switch (fk) {
    case FK_TEST: return launch_bs<FK_TEST>;
    case FK_ABC: return launch_bs<FK_ABC>;
    case FK_ABCV: return launch_bs<FK_ABCV>;
    case FK_GRID_REF: return launch_bs<FK_GRID_REF>;
    case FK_GRID_TB: return launch_bs<FK_GRID_TB>;
    case FK_MODES: return launch_bs<FK_MODES>;
    case FK_WLCS: return launch_bs<FK_WLCS>;
    }
    return nullptr;
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
                         int *nsteps, int *flag, unsigned long long *ctr, int cap)
{
    memset(&a, 0, sizeof(a));
    a.f = c->f;
    bs_consts(p->kmax, p->rtol, &a.c);
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
