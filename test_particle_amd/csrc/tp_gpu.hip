/*
 * tp_gpu.hip -- HIP kernels (gfx950) and the C ABI of include/tp_gpu.h.
 *
 * The reference's hot loop (src/tracking.c:297-462) runs one OpenMP
 * iteration per particle with the time loop inside.  Here one lane owns
 * one particle; a launch advances every particle through a chunk of steps
 * with the state in registers (SoA arrays in HBM between chunks).
 * Diagnostics are fused into the step loop:
 *   - energy spectrum (ptl_energy_fixed, src/diagnostics.c:187-205): an LDS
 *     histogram per block per frame, flushed with 32-bit integer atomics
 *     (exact, order-independent);
 *   - diffusion sums (calc_diff_coeff, src/diagnostics.c:659-705, summed at
 *     src/tracking.c:347-365): a fixed-order wave/LDS reduction per block
 *     into per-block partials, then a fixed-order reduction over blocks
 *     (deterministic run to run; the reference uses omp atomics);
 *   - remaining-particle counts: integer atomics.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <mutex>
#include <new>

#include "../../include/tp_gpu.h"
#include "tp_device.h"

#include "tp_ctx.h"

#include "tp_track.h"

/* ------------------------------------------------------------------ */
/* step kernel: steps j0..j1 of every particle                         */
/* ------------------------------------------------------------------ */
/* One launch covers a run of steps that ends at a diffusion-coefficient
 * step (or at estep); the kernel evaluates the D terms on the state after
 * that step (k_dcoeff does for launches over index lists).  The spectrum
 * (any cadence) and the escape spectrum are binned here. */
/* The kernel's argument block read again from the kernarg segment (the
 * first argument of k_track, k_track_w3 and k_dcoeff, so at offset 0).  Values needed only after the step loop
 * or in its rare branches (output pointers, spectrum settings) are loaded
 * where they are used through this pointer, which the empty asm makes
 * opaque, instead of occupying scalar registers through the loop (which
 * spills them to VGPR lanes and reads them back with a VALU instruction
 * each, every step). */
typedef __attribute__((address_space(4))) const TrackArgs KArgs;
__device__ __forceinline__ KArgs *kargs_again()
{
    KArgs *k = (KArgs *)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(k));
    return k;
}

/* ------------------------------------------------------------------ */
/* diffusion terms at a D step (calc_diff_coeff, src/diagnostics.c:659-705,
 * summed as src/tracking.c:347-369 does)                              */
/* ------------------------------------------------------------------ */
/* Particles inside the box get fresh terms; escaped ones contribute
 * drr = dpp = duu = 0 and their last (stale) dxx dyy dzz daa dee, as the
 * reference's loop-private variables do (src/tracking.c:334-336).  The
 * initial-particle terms gamma_init, mu_init, theta_init were evaluated by
 * k_load with the same operations.  Each block writes one row of 8 partial
 * sums (fixed-order wave and LDS reduction); escaped particles decrement
 * the remaining count of the next row.  The step kernel evaluates them at
 * the end of a launch that ends at a D step (particle g of the launch is
 * particle g, so the blocks and their sums are those of k_dcoeff, which
 * runs instead after launches over an index list). */
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

/* the eight terms of particle i in state (x .. zs) with field e at it (read
 * only when escaped == 0); updates its stale terms */
__device__ __forceinline__ void dterms_particle(KArgs *k, int i, int escaped, double x,
                                                double y, double z, double ux, double uy,
                                                double uz, double tt, double xs, double ys,
                                                double zs, const Emf &e, double v[8])
{
    const int n = k->ntot;
    double *st = k->st;
    double dxx = st[i], dyy = st[n + i], dzz = st[2 * n + i];
    double daa = st[3 * n + i], dee = st[4 * n + i];
    double drr = 0.0, dpp = 0.0, duu = 0.0;
    if (k->fix_stale) dxx = dyy = dzz = daa = dee = 0.0;
    if (escaped == 0) {
        double t = tt * 2 + k->dt;
        double q;
        q = k->x0[i] - (x + xs); dxx = q * q / t;
        q = k->y0[i] - (y + ys); dyy = q * q / t;
        q = k->z0[i] - (z + zs); dzz = q * q / t;
        drr = dxx + dyy + dzz;
        double ax = k->ux0[i] - ux, ay = k->uy0[i] - uy, az = k->uz0[i] - uz;
        dpp = (ax * ax + ay * ay + az * az) / t;
        double g = sqrt(1.0 + ux * ux + uy * uy + uz * uz);
        q = g - k->g0[i]; dee = q * q / t;
        double ab = sqrt(e.bx * e.bx + e.by * e.by + e.bz * e.bz);
        double au = sqrt(ux * ux + uy * uy + uz * uz);
        double mu = (e.bx * ux + e.by * uy + e.bz * uz) / (ab * au);
        q = mu - k->mu0[i]; duu = q * q / t;
        double th = acos(mu) * 180 / TP_PI;
        q = th - k->th0[i]; daa = q * q / t;
    }
    st[i] = dxx; st[n + i] = dyy; st[2 * n + i] = dzz;
    st[3 * n + i] = daa; st[4 * n + i] = dee;
    v[0] = drr; v[1] = dxx; v[2] = dyy; v[3] = dzz;
    v[4] = dpp; v[5] = duu; v[6] = daa; v[7] = dee;
}

/* the block's row of partial sums and its escape count */
__device__ __forceinline__ void dterms_reduce(KArgs *k, const double v[8], bool live,
                                              int escaped, int tid, int tstep,
                                              double (*wsum)[8])
{
    const int lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        double r = wave_sum(v[q]);
        if (lane == 0) wsum[wave][q] = r;
    }
    /* the escape count: one atomic per block (per-wave atomics on the one
     * counter serialise when every lane is escaped) */
    __shared__ int wesc[4];
    const unsigned long long m = __ballot(live && escaped);
    if (lane == 0) wesc[wave] = __popcll(m);
    __syncthreads();
    if (tid < 8) {
        double r = 0.0;
        for (int w = 0; w < nwaves; w++) r += wsum[w][tid];
        k->partials[(size_t)(blockIdx.x + k->pboff) * 8 + tid] = r;
    }
    if (tid == 0 && tstep + 1 < k->nd) {
        int e = 0;
        for (int w = 0; w < nwaves; w++) e += wesc[w];
        if (e) atomicAdd(&k->rdec[tstep + 1], e);
    }
}

/* The particle's position shifts (xs, ys, zs) are read only when a wrap
 * first changes them, or when the D terms or a trajectory record need
 * them, and written back only when a wrap changed them: most particles
 * cross no box face during a launch, so this saves 48 of the 176 bytes of
 * state a launch reads and writes per particle.  The wraps themselves are
 * applied to the loaded values in the same order (bit-identical shifts). */
__device__ __forceinline__ void load_shifts(int i, double &xs, double &ys, double &zs, bool &have)
{
    if (!have) {
        KArgs *k = kargs_again();
        xs = k->xs[i]; ys = k->ys[i]; zs = k->zs[i];
        have = true;
    }
}

/* particle_bc (src/tracking.c:1172-1221) with the shifts read on the first
 * wrap */
__device__ __forceinline__ void particle_bc_lazy(const TrackArgs &a, Ptl &p, int i, double &xs,
                                                 double &ys, double &zs, bool &have,
                                                 bool &wrapped, int &esc)
{
    if (a.bc == 0) {
        if (p.x > a.hi[0] || p.y > a.hi[1] || p.z > a.hi[2] || p.x < a.lo[0] ||
            p.y < a.lo[1] || p.z < a.lo[2]) {
            load_shifts(i, xs, ys, zs, have);
            wrapped = true;
            particle_bc(a, p, xs, ys, zs, esc);
        }
    } else {
        particle_bc(a, p, xs, ys, zs, esc);
    }
}

template <int PUSH, int FK, int PREC, bool TRAJ>
__device__ __forceinline__ void track_body(const TrackArgs &a)
{
    extern __shared__ unsigned hist[];
    const int tid = threadIdx.x;
    const int gi = blockIdx.x * blockDim.x + tid;
    const bool live = gi < a.n;
    const int i = live ? (a.idx ? a.idx[gi] : gi) : 0;
    for (int b = tid; b < a.nbins; b += blockDim.x) hist[b] = 0u;

    Ptl p;
    double xs = 0.0, ys = 0.0, zs = 0.0;
    bool have_shift = false, wrapped = false;
    int esc = 1, esc0 = 1, nst = 0;
    if (live) {
        p.x = a.x[i]; p.y = a.y[i]; p.z = a.z[i];
        p.ux = a.ux[i]; p.uy = a.uy[i]; p.uz = a.uz[i]; p.t = a.t[i];
        if (TRAJ) load_shifts(i, xs, ys, zs, have_shift);
        esc = a.esc[i];
        esc0 = esc;
        nst = a.nsteps[i];
        if (a.first) {
            /* initial wrap and escape check, src/tracking.c:314-332 */
            esc = 0;
            nst = 1;
            particle_bc_lazy(a, p, i, xs, ys, zs, have_shift, wrapped, esc);
            if (esc && 1 < a.nd) atomicAdd(&a.rdec[1], 1);
        }
    } else {
        p.x = p.y = p.z = p.ux = p.uy = p.uz = p.t = 0.0;
    }
    float fux = (float)p.ux, fuy = (float)p.uy, fuz = (float)p.uz;
    const float dtf = (float)a.dt;
    const float hf = (float)(a.qm * (0.5 * a.dt));
    /* trajectory diagnostics: this particle's next slot (its slot range is
     * read when a point is recorded) */
    int tpts = (TRAJ && live) ? a.traj_pts[i] : 0;
    typename CacheFor<FK>::type gc;
    gc_init(gc);
    __syncthreads();

    int ntp = a.ntp0;
    for (int j = a.j0; j <= a.j1; j++) {
        int esc_after = esc;
        if (live && (esc == 0 || a.bc == 0)) {
            if (PREC == PREC64) {
                if (PUSH == TP_PUSH_WIRZ) wirz_step64<FK>(a.f, gc, p, a.dt, a.qm);
                else boris_step64<FK>(a.f, gc, p, a.dt, a.qm);
            } else {
                boris_step32<FK>(a.f, gc, p, fux, fuy, fuz, a.dt, dtf, hf);
                p.ux = fux; p.uy = fuy; p.uz = fuz;
            }
            particle_bc_lazy(a, p, i, xs, ys, zs, have_shift, wrapped, esc_after);
            nst++;
            /* src/tracking.c:440-447: the particle after the step, every
             * nsteps_output steps; slots past the particle's range are
             * counted, not written (the reference writes them) */
            if (TRAJ && j % a.traj_nout == 0) {
                const int toff = i > 0 ? a.traj_accum[i - 1] : 0;
                if (tpts < a.traj_accum[i] - toff) {
                    tp_particle r;
                    r.x = p.x; r.y = p.y; r.z = p.z; r.qm = a.stage[i].qm;
                    r.ux = p.ux; r.uy = p.uy; r.uz = p.uz; r.t = p.t;
                    r.xshift = xs; r.yshift = ys; r.zshift = zs;
                    a.traj[toff + tpts] = r;
                }
                tpts++;
            }
        }
        if (live && esc == 0 && esc_after == 1 && a.bc == 1) {
            KArgs *k = kargs_again();
            if (ntp < k->nt_out)
                atomicAdd(&k->espect_esc[ntp * k->nbins +
                          energy_bin(p.ux, p.uy, p.uz, k->nbins, k->logemin, k->logde)], 1u);
        }
        if (j % a.einterval == 0) {
            KArgs *k = kargs_again();
            const int nt_out = k->nt_out, nbins = k->nbins;
            if (live && esc_after == 0 && ntp < nt_out)
                atomicAdd(&hist[energy_bin(p.ux, p.uy, p.uz, nbins, k->logemin, k->logde)], 1u);
            __syncthreads();
            if (ntp < nt_out) {
                unsigned *espect = k->espect;
                for (int b = tid; b < nbins; b += blockDim.x) {
                    unsigned h = hist[b];
                    if (h) {
                        atomicAdd(&espect[ntp * nbins + b], h);
                        hist[b] = 0u;
                    }
                }
            }
            __syncthreads();
            ntp++;
        }
        esc = esc_after;
    }
    if (a.dstep >= 0) {
        /* the D terms of the state after step j1 (a D step) */
        __shared__ double wsum[4][8];
        KArgs *k = kargs_again();
        double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (live) {
            load_shifts(i, xs, ys, zs, have_shift);
            Emf e = {};
            if (esc == 0) {
                if constexpr (PREC == PREC64)
                    emf64_step<FK>(a.f, gc, p.x, p.y, p.z, p.t, e);
                else if constexpr (FK == FK_GRID_REF || FK == FK_GRID2_REF)
                    ref_emf64_via_cache<FK>(a.f, gc, p.x, p.y, p.z, p.t, e);
                else if constexpr (FK == FK_MHD_REF) {
                    ref_emf64_via_cache<FK_GRID_REF>(a.f, gc, p.x, p.y, p.z, p.t, e);
                    mhd_e64(e);
                }
                else
                    get_emf64<FK>(a.f, p.x, p.y, p.z, p.t, e);
            }
            dterms_particle(k, i, esc, p.x, p.y, p.z, p.ux, p.uy, p.uz, p.t, xs, ys, zs, e, v);
        }
        dterms_reduce(k, v, live, live ? esc : 0, tid, a.dstep, wsum);
    }
    if (live) {
        KArgs *k = kargs_again();
        k->x[i] = p.x; k->y[i] = p.y; k->z[i] = p.z;
        k->ux[i] = p.ux; k->uy[i] = p.uy; k->uz[i] = p.uz; k->t[i] = p.t;
        if (wrapped) { k->xs[i] = xs; k->ys[i] = ys; k->zs[i] = zs; }
        if (esc != esc0) k->esc[i] = esc;
        k->nsteps[i] = nst;
        if (TRAJ) k->traj_pts[i] = tpts;
    }
}

template <int PUSH, int FK, int PREC, bool TRAJ>
__global__ void __launch_bounds__(256) k_track(TrackArgs a)
{
    track_body<PUSH, FK, PREC, TRAJ>(a);
}

/* the same compiled for 3 waves per SIMD (168 VGPRs; the compiler's own
 * choice is 2 for the kernels launched this way) */
template <int PUSH, int FK, int PREC, bool TRAJ>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_track_w3(TrackArgs a)
{
    track_body<PUSH, FK, PREC, TRAJ>(a);
}

template <int FK>
__global__ void __launch_bounds__(256) k_dcoeff(TrackArgs a, int tstep)
{
    __shared__ double wsum[4][8];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * blockDim.x + tid;
    const bool live = i < a.n;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int escaped = 0;
    KArgs *k = kargs_again();
    if (live) {
        escaped = a.esc[i];
        Emf e = {};
        double x = 0.0, y = 0.0, z = 0.0, ux = 0.0, uy = 0.0, uz = 0.0, tt = 0.0;
        double xs = 0.0, ys = 0.0, zs = 0.0;
        if (escaped == 0) {
            x = a.x[i]; y = a.y[i]; z = a.z[i];
            ux = a.ux[i]; uy = a.uy[i]; uz = a.uz[i]; tt = a.t[i];
            xs = a.xs[i]; ys = a.ys[i]; zs = a.zs[i];
            get_emf64<FK>(a.f, x, y, z, tt, e);
        }
        dterms_particle(k, i, escaped, x, y, z, ux, uy, uz, tt, xs, ys, zs, e, v);
    }
    dterms_reduce(k, v, live, escaped, tid, tstep, wsum);
}

/* The D terms of escaped particles listed in idx[0 .. n): their stale
 * terms (and the escape count), as the step kernel forms them for its
 * escaped lanes; block rows from pboff on. */
__global__ void __launch_bounds__(256) k_dstale(TrackArgs a, int tstep)
{
    __shared__ double wsum[4][8];
    const int tid = threadIdx.x;
    const int g = blockIdx.x * blockDim.x + tid;
    const bool live = g < a.n;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    KArgs *k = kargs_again();
    if (live) {
        const Emf e = {};
        dterms_particle(k, a.idx[g], 1, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, e, v);
    }
    dterms_reduce(k, v, live, live ? 1 : 0, tid, tstep, wsum);
}

/* Open boundaries: stable split of the particles into those still inside
 * the box (esc == 0) and those that left, each list in index order, in
 * three passes over blocks of 256: counts, one-block scan, scatter. */
__global__ void __launch_bounds__(256) k_live_count(const int *esc, int n, int *cnt)
{
    __shared__ int w[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const unsigned long long m = __ballot(i < n && esc[i] == 0);
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = w[0] + w[1] + w[2] + w[3];
}

__global__ void __launch_bounds__(1024) k_live_scan(const int *cnt, int nb, int *off,
                                                    int *total)
{
    __shared__ int s[1024];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int b0 = t * per, b1 = min(nb, b0 + per);
    int sum = 0;
    for (int b = b0; b < b1; b++) sum += cnt[b];
    s[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {          /* inclusive scan of the chunk sums */
        const int v = t >= o ? s[t - o] : 0;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    int run = s[t] - sum;
    for (int b = b0; b < b1; b++) {
        off[b] = run;
        run += cnt[b];
    }
    if (t == 1023) *total = s[1023];
}

__global__ void __launch_bounds__(256) k_live_scatter(const int *esc, int n, const int *off,
                                                      int *inside, int *gone)
{
    __shared__ int w[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < n;
    const bool in = valid && esc[i] == 0;
    const unsigned long long m = __ballot(in);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) w[wave] = __popcll(m);
    __syncthreads();
    int rl = __popcll(m & ((1ull << lane) - 1ull));   /* inside, before i in the block */
    for (int q = 0; q < wave; q++) rl += w[q];
    const int base = off[blockIdx.x];                /* inside, in earlier blocks */
    if (in) inside[base + rl] = i;
    else if (valid) gone[(blockIdx.x * 256 - base) + ((int)threadIdx.x - rl)] = i;
}

/* Sum the per-block partials of one D row in a fixed order: one block per
 * component, strided per-thread sums then an LDS tree. */
__global__ void __launch_bounds__(256) k_reduce_row(const double *partials, int nblocks,
                                                    int row, int nd, double *dco)
{
    __shared__ double sh[256];
    const int k = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256) s += partials[(size_t)b * 8 + k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0 && row < nd) dco[(size_t)k * nd + row] = sh[0];
}

/* AoS (struct particles) -> SoA, for both the current and the initial
 * particles; the initial ones also get gamma_init, mu_init and theta_init
 * as calc_diff_coeff evaluates them (src/diagnostics.c:680-703). */
template <int FK>
__global__ void __launch_bounds__(256) k_load(FieldDev f, int n, const tp_particle *cur,
        const tp_particle *ini, double *x, double *y, double *z, double *ux,
        double *uy, double *uz, double *t, double *xs, double *ys, double *zs,
        double *x0, double *y0, double *z0, double *ux0, double *uy0, double *uz0,
        double *t0, double *g0, double *mu0, double *th0, double *st, int *esc,
        int *nsteps)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tp_particle c = cur[i], p = ini[i];
    x[i] = c.x; y[i] = c.y; z[i] = c.z; ux[i] = c.ux; uy[i] = c.uy; uz[i] = c.uz;
    t[i] = c.t; xs[i] = c.xshift; ys[i] = c.yshift; zs[i] = c.zshift;
    x0[i] = p.x; y0[i] = p.y; z0[i] = p.z; ux0[i] = p.ux; uy0[i] = p.uy; uz0[i] = p.uz;
    t0[i] = p.t;
    g0[i] = sqrt(1.0 + p.ux * p.ux + p.uy * p.uy + p.uz * p.uz);
    Emf e;
    get_emf64<FK>(f, p.x, p.y, p.z, p.t, e);
    double ab = sqrt(e.bx * e.bx + e.by * e.by + e.bz * e.bz);
    double au = sqrt(p.ux * p.ux + p.uy * p.uy + p.uz * p.uz);
    double m0 = (e.bx * p.ux + e.by * p.uy + e.bz * p.uz) / (ab * au);
    mu0[i] = m0;
    th0[i] = acos(m0) * 180 / TP_PI;
    for (int k = 0; k < 5; k++) st[(size_t)k * n + i] = 0.0;
    esc[i] = 0;
    nsteps[i] = 1;
}

__global__ void __launch_bounds__(256) k_store(int n, const double *x, const double *y,
        const double *z, const double *ux, const double *uy, const double *uz,
        const double *t, const double *xs, const double *ys, const double *zs,
        tp_particle *out)
{
    /* out[] still holds the particles as uploaded; qm is not tracked */
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tp_particle c;
    c.x = x[i]; c.y = y[i]; c.z = z[i]; c.qm = out[i].qm;
    c.ux = ux[i]; c.uy = uy[i]; c.uz = uz[i]; c.t = t[i];
    c.xshift = xs[i]; c.yshift = ys[i]; c.zshift = zs[i];
    out[i] = c;
}

template <int FK, int PREC>
__global__ void __launch_bounds__(256) k_eval(FieldDev f, int n, const double *pts,
                                              double *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = pts[4 * i], y = pts[4 * i + 1], z = pts[4 * i + 2], t = pts[4 * i + 3];
    double *o = out + 6 * i;
    if (PREC == PREC64) {
        Emf e;
        get_emf64<FK>(f, x, y, z, t, e);
        o[0] = e.bx; o[1] = e.by; o[2] = e.bz; o[3] = e.ex; o[4] = e.ey; o[5] = e.ez;
    } else {
        EmfF e;
        get_emf32<FK>(f, x, y, z, t, e);
        o[0] = e.bx; o[1] = e.by; o[2] = e.bz; o[3] = e.ex; o[4] = e.ey; o[5] = e.ez;
    }
}

/* ------------------------------------------------------------------ */
/* dispatch                                                            */
/* ------------------------------------------------------------------ */



template <int PUSH, int FK, int PREC, bool TRAJ>
static void launch_track(dim3 g, dim3 b, size_t sh, hipStream_t s, const TrackArgs &a)
{
    /* 3 waves per SIMD: the FP32 two-slice gather (no spills on its hot
     * path) and the Wirz push through the reference gather (a few spills;
     * measured 4 % faster than the compiler's 2 waves) */
    if constexpr ((PREC == PREC32 && (FK == FK_GRID2_REF || FK == FK_GRID2_TB)) ||
                  (PUSH == TP_PUSH_WIRZ && (FK == FK_GRID_REF || FK == FK_MHD_REF)))
        hipLaunchKernelGGL((k_track_w3<PUSH, FK, PREC, TRAJ>), g, b, sh, s, a);
    else
        hipLaunchKernelGGL((k_track<PUSH, FK, PREC, TRAJ>), g, b, sh, s, a);
}

/* trajectory recording is a separate instantiation, so the kernels without
 * it carry none of its registers */
template <int FK, bool TRAJ>
static track_fn pick_track_fkt(int push, int prec)
{
    if (prec == TP_FP64) {
        if (push == TP_PUSH_WIRZ) return launch_track<TP_PUSH_WIRZ, FK, PREC64, TRAJ>;
        return launch_track<TP_PUSH_BORIS, FK, PREC64, TRAJ>;
    }
    if (push == TP_PUSH_BORIS) return launch_track<TP_PUSH_BORIS, FK, PREC32, TRAJ>;
    return nullptr;
}

template <int FK>
static track_fn pick_track_fk(int push, int prec, bool traj)
{
    return traj ? pick_track_fkt<FK, true>(push, prec) : pick_track_fkt<FK, false>(push, prec);
}

static track_fn pick_track(int fk, int push, int prec, bool traj)
{
    switch (fk) {
    case FK_TEST: return pick_track_fk<FK_TEST>(push, prec, traj);
    case FK_ABC: return pick_track_fk<FK_ABC>(push, prec, traj);
    case FK_ABCV: return pick_track_fk<FK_ABCV>(push, prec, traj);
    case FK_GRID_REF: return pick_track_fk<FK_GRID_REF>(push, prec, traj);
    case FK_GRID_TB: return pick_track_fk<FK_GRID_TB>(push, prec, traj);
    case FK_MODES: return pick_track_fk<FK_MODES>(push, prec, traj);
    case FK_WLCS: return pick_track_fk<FK_WLCS>(push, prec, traj);
    case FK_GRID2_REF: return pick_track_fk<FK_GRID2_REF>(push, prec, traj);
    case FK_GRID2_TB: return pick_track_fk<FK_GRID2_TB>(push, prec, traj);
    case FK_MHD_REF: return pick_track_fk<FK_MHD_REF>(push, prec, traj);
    case FK_GRIDD_REF: return pick_track_fk<FK_GRIDD_REF>(push, prec, traj);
    }
    return nullptr;
}

template <int FK>
static void launch_dcoeff(dim3 g, dim3 b, hipStream_t s, const TrackArgs &a, int tstep)
{
    hipLaunchKernelGGL((k_dcoeff<FK>), g, b, 0, s, a, tstep);
}

static void dispatch_dcoeff(int fk, dim3 g, dim3 b, hipStream_t s, const TrackArgs &a,
                            int tstep)
{
    switch (fk) {
    case FK_TEST: launch_dcoeff<FK_TEST>(g, b, s, a, tstep); break;
    case FK_ABC: launch_dcoeff<FK_ABC>(g, b, s, a, tstep); break;
    case FK_ABCV: launch_dcoeff<FK_ABCV>(g, b, s, a, tstep); break;
    case FK_GRID_REF: launch_dcoeff<FK_GRID_REF>(g, b, s, a, tstep); break;
    case FK_GRID_TB: launch_dcoeff<FK_GRID_TB>(g, b, s, a, tstep); break;
    case FK_MODES: launch_dcoeff<FK_MODES>(g, b, s, a, tstep); break;
    case FK_WLCS: launch_dcoeff<FK_WLCS>(g, b, s, a, tstep); break;
    case FK_MLP: launch_dcoeff<FK_MODES>(g, b, s, a, tstep); break;
    case FK_GRID2_REF: launch_dcoeff<FK_GRID2_REF>(g, b, s, a, tstep); break;
    case FK_GRID2_TB: launch_dcoeff<FK_GRID2_TB>(g, b, s, a, tstep); break;
    case FK_MHD_REF: launch_dcoeff<FK_MHD_REF>(g, b, s, a, tstep); break;
    case FK_GRIDD_REF: launch_dcoeff<FK_GRIDD_REF>(g, b, s, a, tstep); break;
    }
}

template <int FK>
static void launch_load(dim3 g, dim3 b, hipStream_t s, const FieldDev &f, int n,
                        const tp_particle *cur, const tp_particle *ini, State &S)
{
    hipLaunchKernelGGL((k_load<FK>), g, b, 0, s, f, n, cur, ini, S.x, S.y, S.z, S.ux,
                       S.uy, S.uz, S.t, S.xs, S.ys, S.zs, S.x0, S.y0, S.z0, S.ux0,
                       S.uy0, S.uz0, S.t0, S.g0, S.mu0, S.th0, S.st, S.esc, S.nsteps);
}

static void dispatch_load(int fk, dim3 g, dim3 b, hipStream_t s, const FieldDev &f,
                          int n, const tp_particle *cur, const tp_particle *ini,
                          State &S)
{
    switch (fk) {
    case FK_TEST: launch_load<FK_TEST>(g, b, s, f, n, cur, ini, S); break;
    case FK_ABC: launch_load<FK_ABC>(g, b, s, f, n, cur, ini, S); break;
    case FK_ABCV: launch_load<FK_ABCV>(g, b, s, f, n, cur, ini, S); break;
    case FK_GRID_REF: launch_load<FK_GRID_REF>(g, b, s, f, n, cur, ini, S); break;
    case FK_GRID_TB: launch_load<FK_GRID_TB>(g, b, s, f, n, cur, ini, S); break;
    case FK_MODES: launch_load<FK_MODES>(g, b, s, f, n, cur, ini, S); break;
    case FK_WLCS: launch_load<FK_WLCS>(g, b, s, f, n, cur, ini, S); break;
    case FK_MLP: launch_load<FK_MODES>(g, b, s, f, n, cur, ini, S); break;
    case FK_GRID2_REF: launch_load<FK_GRID2_REF>(g, b, s, f, n, cur, ini, S); break;
    case FK_GRID2_TB: launch_load<FK_GRID2_TB>(g, b, s, f, n, cur, ini, S); break;
    case FK_MHD_REF: launch_load<FK_MHD_REF>(g, b, s, f, n, cur, ini, S); break;
    case FK_GRIDD_REF: launch_load<FK_GRIDD_REF>(g, b, s, f, n, cur, ini, S); break;
    }
}

template <int FK>
static void launch_eval(dim3 g, dim3 b, hipStream_t s, const FieldDev &f, int n,
                        const double *pts, double *out, int prec)
{
    if (prec == TP_FP64) hipLaunchKernelGGL((k_eval<FK, PREC64>), g, b, 0, s, f, n, pts, out);
    else hipLaunchKernelGGL((k_eval<FK, PREC32>), g, b, 0, s, f, n, pts, out);
}

static void dispatch_eval(int fk, dim3 g, dim3 b, hipStream_t s, const FieldDev &f,
                          int n, const double *pts, double *out, int prec)
{
    switch (fk) {
    case FK_TEST: launch_eval<FK_TEST>(g, b, s, f, n, pts, out, prec); break;
    case FK_ABC: launch_eval<FK_ABC>(g, b, s, f, n, pts, out, prec); break;
    case FK_ABCV: launch_eval<FK_ABCV>(g, b, s, f, n, pts, out, prec); break;
    case FK_GRID_REF: launch_eval<FK_GRID_REF>(g, b, s, f, n, pts, out, prec); break;
    case FK_GRID_TB: launch_eval<FK_GRID_TB>(g, b, s, f, n, pts, out, prec); break;
    case FK_MODES: launch_eval<FK_MODES>(g, b, s, f, n, pts, out, prec); break;
    case FK_WLCS: launch_eval<FK_WLCS>(g, b, s, f, n, pts, out, prec); break;
    case FK_MLP: launch_eval<FK_MODES>(g, b, s, f, n, pts, out, prec); break;
    case FK_GRID2_REF: launch_eval<FK_GRID2_REF>(g, b, s, f, n, pts, out, prec); break;
    case FK_GRID2_TB: launch_eval<FK_GRID2_TB>(g, b, s, f, n, pts, out, prec); break;
    case FK_MHD_REF: launch_eval<FK_MHD_REF>(g, b, s, f, n, pts, out, prec); break;
    case FK_GRIDD_REF: launch_eval<FK_GRIDD_REF>(g, b, s, f, n, pts, out, prec); break;
    }
}

/* ------------------------------------------------------------------ */
/* C ABI                                                               */
/* ------------------------------------------------------------------ */
extern "C" {

const char *tp_gpu_strerror(int code)
{
    switch (code) {
    case TP_OK: return "ok";
    case TP_E_ARG: return "invalid argument";
    case TP_E_NOMEM: return "out of memory";
    case TP_E_HIP: return "HIP runtime error";
    case TP_E_NONFINITE: return "non-finite particle state";
    case TP_E_STATE: return "call out of order";
    case TP_E_NODEV: return "no usable GPU";
    }
    return "unknown error";
}

static void *noalloc_scratch(int dev);

int tp_gpu_init(int device, tp_ctx **out)
{
    if (!out) return TP_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TP_E_NODEV;
    if (device < 0 || device >= ndev) return TP_E_ARG;
    tp_ctx *c = new (std::nothrow) tp_ctx();
    if (!c) return TP_E_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return TP_E_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
    memset(&c->f, 0, sizeof(c->f));
    /* the scratch of the allocation-free run, looked up with the device current */
    c->scratch = (double *)noalloc_scratch(device);
    *out = c;
    return TP_OK;
}

static void free_field(tp_ctx *c)
{
    (void)hipFree(c->grid); c->grid = nullptr;
    (void)hipFree(c->grid_b); c->grid_b = nullptr;
    (void)hipFree(c->cells); c->cells = nullptr;
    (void)hipFree(c->cells_b); c->cells_b = nullptr;
    (void)hipFree(c->vgrid); c->vgrid = nullptr;
    (void)hipFree(c->modes); c->modes = nullptr;
    (void)hipFree(c->modes32); c->modes32 = nullptr;
    (void)hipFree(c->wl); c->wl = nullptr;
    (void)hipFree(c->gridd); c->gridd = nullptr;
    c->have_field = false;
    c->s.derived = false;
}

namespace {
void staging_release(tp_ctx *c);
}

void tp_gpu_finalize(tp_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_field(c);
    (void)hipFree(c->s.buf);
    (void)hipFree(c->s.ibuf);
    (void)hipFree(c->s.stage);
    (void)hipFree(c->espect);
    (void)hipFree(c->espect_esc);
    (void)hipFree(c->dco);
    (void)hipFree(c->partials);
    (void)hipFree(c->rdec);
    (void)hipFree(c->live_idx);
    for (int k = 0; k < c->nev; k++) (void)hipEventDestroy(c->evs[k]);
    free(c->evs);
    tp_bs_release(c);
    tp_mlp_release(c);
    (void)hipFree(c->traj);
    (void)hipFree(c->traj_accum);
    (void)hipFree(c->traj_pts);
    free(c->bs_tune_block);
    staging_release(c);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int tp_gpu_device_info(tp_ctx *c, char *name, int name_len, int *num_cu)
{
    if (!c) return TP_E_ARG;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, c->device));
    if (name && name_len > 0) {
        snprintf(name, name_len, "%s%s(%s)", prop.name, prop.name[0] ? " " : "",
                 prop.gcnArchName);
    }
    if (num_cu) *num_cu = prop.multiProcessorCount;
    return TP_OK;
}

static void set_box(double *dst_lo, double *dst_hi, double *dst_d, const double *lo,
                    const double *hi, int nx, int ny, int nz)
{
    int n[3] = {nx, ny, nz};
    for (int a = 0; a < 3; a++) {
        dst_lo[a] = lo[a];
        dst_hi[a] = hi[a];
        dst_d[a] = (hi[a] - lo[a]) / (n[a] - 1.0);   /* src/domain.c:548-550 */
    }
}

/* 1 / d per axis, and whether every d is a power of two with a normal
 * reciprocal: then v / d and v * (1 / d) round the same real number, so the
 * FP64 gathers may multiply */
static void set_inv_d(FieldDev &f)
{
    f.exact_inv = 1;
    for (int a = 0; a < 3; a++) {
        f.inv_d[a] = 1.0 / f.d[a];
        int e = 0;
        const double m = frexp(f.d[a], &e);
        if (!(m == 0.5 && isnormal(f.d[a]) && isnormal(f.inv_d[a]))) f.exact_inv = 0;
    }
}

int tp_gpu_set_field_test(tp_ctx *c)
{
    if (!c) return TP_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    free_field(c);
    memset(&c->f, 0, sizeof(c->f));
    c->f.kind = FK_TEST;
    c->have_field = true;
    return TP_OK;
}

int tp_gpu_set_field_abc(tp_ctx *c, double A, double B, double C, double lambda, double B0)
{
    if (!c) return TP_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    /* keep a velocity grid set before */
    double *vg = c->vgrid;
    c->vgrid = nullptr;
    FieldDev keep = c->f;
    free_field(c);
    c->vgrid = vg;
    memset(&c->f, 0, sizeof(c->f));
    c->f.A = A; c->f.B = B; c->f.C = C; c->f.lam = lambda; c->f.B0 = B0;
    if (vg) {
        c->f.vnx = keep.vnx; c->f.vny = keep.vny; c->f.vnz = keep.vnz;
        memcpy(c->f.vlo, keep.vlo, sizeof(keep.vlo));
        memcpy(c->f.vhi, keep.vhi, sizeof(keep.vhi));
        memcpy(c->f.vd, keep.vd, sizeof(keep.vd));
        c->f.vgrid = vg;
        c->f.kind = FK_ABCV;
    } else {
        c->f.kind = FK_ABC;
    }
    c->have_field = true;
    return TP_OK;
}

int tp_gpu_set_velocity_grid(tp_ctx *c, int nx, int ny, int nz, const double lo[3],
                             const double hi[3], const double *v)
{
    if (!c) return TP_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    (void)hipFree(c->vgrid);
    c->vgrid = nullptr;
    c->f.vgrid = nullptr;
    if (c->f.kind == FK_ABCV) c->f.kind = FK_ABC;
    c->s.derived = false;
    if (!v) return TP_OK;
    if (nx < 2 || ny < 2 || nz < 2 || !lo || !hi) return TP_E_ARG;
    if ((size_t)nx * ny * nz >= 0xFFFFFFFFull) return TP_E_ARG;   /* 32-bit corner indices */
    size_t bytes = (size_t)nx * ny * nz * 3 * sizeof(double);
    if (hipMalloc(&c->vgrid, bytes) != hipSuccess) { c->vgrid = nullptr; return TP_E_NOMEM; }
    HIPCHK(hipMemcpy(c->vgrid, v, bytes, hipMemcpyHostToDevice));
    c->f.vnx = nx; c->f.vny = ny; c->f.vnz = nz;
    set_box(c->f.vlo, c->f.vhi, c->f.vd, lo, hi, nx, ny, nz);
    c->f.vgrid = c->vgrid;
    if (c->f.kind == FK_ABC) c->f.kind = FK_ABCV;
    return TP_OK;
}

/* ------------------------------------------------------------------ */
/* PIC grid input: raw component planes or the reference's float3 node */
/* arrays -> 32-byte nodes {ex ey ez bx | by bz 0 0}, x fastest, built */
/* on the GPU from slabs staged through pinned host memory             */
/* ------------------------------------------------------------------ */

/* the reference's scaling: data[i] * E0 in double, stored as float
 * (src/electric_field.c:276, :481) */
__device__ __forceinline__ float plane_scale(float v, double s) { return (float)((double)v * s); }

__device__ __forceinline__ void store_node(float4 *grid, size_t node, const float v[6],
                                           double e0, double b0)
{
    grid[2 * node] = make_float4(plane_scale(v[0], e0), plane_scale(v[1], e0),
                                 plane_scale(v[2], e0), plane_scale(v[3], b0));
    grid[2 * node + 1] = make_float4(plane_scale(v[4], b0), plane_scale(v[5], b0), 0.f, 0.f);
}

/* x-fastest planes: a slab of cnt nodes starting at node0; plane c of the
 * slab at src + c*cnt.  Reads and writes are both contiguous in the node
 * index, so one lane per node. */
__global__ void __launch_bounds__(256) k_planes_xf(const float *__restrict__ src, size_t cnt,
                                                   size_t node0, float4 *__restrict__ grid,
                                                   double e0, double b0)
{
    for (size_t n = (size_t)blockIdx.x * 256 + threadIdx.x; n < cnt;
         n += (size_t)gridDim.x * 256) {
        float v[6];
#pragma unroll
        for (int c = 0; c < 6; c++) v[c] = src[c * cnt + n];
        store_node(grid, node0 + n, v, e0, b0);
    }
}

/* x-slowest planes: a slab of ci x-indices starting at i0, plane c of the
 * slab at src + c*ci*ny*nz with index (il*ny + j)*nz + k.  A block moves a
 * 32 (k) x 32 (il) tile of one j through LDS: the loads run along k, the node
 * stores along x (32 consecutive nodes = 1 KiB per row). */
__global__ void __launch_bounds__(256) k_planes_xs(const float *__restrict__ src, int ci, int i0,
                                                   int nx, int ny, int nz,
                                                   float4 *__restrict__ grid, double e0,
                                                   double b0)
{
    __shared__ float t[6][32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int k0 = blockIdx.x * 32, il0 = blockIdx.y * 32;
    const size_t cnt = (size_t)ci * ny * nz;
    for (int j = blockIdx.z; j < ny; j += gridDim.z) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int il = il0 + ty + 8 * r, k = k0 + tx;
            if (il < ci && k < nz) {
                const size_t at = ((size_t)il * ny + j) * nz + k;
#pragma unroll
                for (int c = 0; c < 6; c++) t[c][ty + 8 * r][tx] = src[c * cnt + at];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int k = k0 + ty + 8 * r, il = il0 + tx;
            if (il < ci && k < nz) {
                float v[6];
#pragma unroll
                for (int c = 0; c < 6; c++) v[c] = t[c][tx][ty + 8 * r];
                store_node(grid, ((size_t)k * ny + j) * nx + i0 + il, v, e0, b0);
            }
        }
        __syncthreads();
    }
}

/* float3 E and B node arrays (x fastest, already scaled): a slab of cnt
 * nodes, E at src, B at src + 3*cnt. */
__global__ void __launch_bounds__(256) k_nodes_aos(const float *__restrict__ src, size_t cnt,
                                                  size_t node0, float4 *__restrict__ grid)
{
    for (size_t n = (size_t)blockIdx.x * 256 + threadIdx.x; n < cnt;
         n += (size_t)gridDim.x * 256) {
        const float *e = src + 3 * n, *b = src + 3 * cnt + 3 * n;
        grid[2 * (node0 + n)] = make_float4(e[0], e[1], e[2], b[0]);
        grid[2 * (node0 + n) + 1] = make_float4(b[1], b[2], 0.f, 0.f);
    }
}

/* source layouts of the pipeline: the two public plane orders, and the
 * reference's float3 node arrays (tp_gpu_set_field_grid / _grid2), read
 * as component 0 (E, 3 floats per node) and 1 (B) */
enum { LAYOUT_XF = TP_PLANES_X_FASTEST, LAYOUT_XS = TP_PLANES_X_SLOWEST, LAYOUT_AOS = 2 };

/* cell (i, j, k) of the cell-major copy: the eight corners of the reference
 * gather's cell with lower corner (i, j, k), in its corner order (0: x1 y1 z1,
 * 1: x1 y1 z2, 2: x1 y2 z1, 3: x2 y1 z1, 4: x2 y2 z1, 5: x2 y1 z2,
 * 6: x1 y2 z2, 7: x2 y2 z2), upper corner i + 1 or 0 after the last node;
 * six floats per corner, written as 12 float4. */
__global__ void __launch_bounds__(256) k_build_cells(const float4 *__restrict__ grid, int nx,
                                                     int ny, int nz, float4 *__restrict__ cells)
{
    const size_t nn = (size_t)nx * ny * nz;
    for (size_t cid = (size_t)blockIdx.x * 256 + threadIdx.x; cid < nn;
         cid += (size_t)gridDim.x * 256) {
        const unsigned i = (unsigned)(cid % (unsigned)nx);
        const unsigned j = (unsigned)((cid / (unsigned)nx) % (unsigned)ny);
        const unsigned k = (unsigned)(cid / ((size_t)nx * ny));
        const unsigned X1 = i, X2 = i + 1 < (unsigned)nx ? i + 1 : 0u;
        const unsigned sx = (unsigned)nx, sxy = (unsigned)nx * (unsigned)ny;
        const unsigned Y1 = sx * j, Y2 = sx * (j + 1 < (unsigned)ny ? j + 1 : 0u);
        const unsigned Z1 = sxy * k, Z2 = sxy * (k + 1 < (unsigned)nz ? k + 1 : 0u);
        const unsigned c[8] = {X1 + Y1 + Z1, X1 + Y1 + Z2, X1 + Y2 + Z1, X2 + Y1 + Z1,
                               X2 + Y2 + Z1, X2 + Y1 + Z2, X1 + Y2 + Z2, X2 + Y2 + Z2};
        float v[48];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const float4 a = grid[2 * (size_t)c[q]];
            const float4 b = grid[2 * (size_t)c[q] + 1];
            v[6 * q + 0] = a.x; v[6 * q + 1] = a.y; v[6 * q + 2] = a.z;
            v[6 * q + 3] = a.w; v[6 * q + 4] = b.x; v[6 * q + 5] = b.y;
        }
        float4 *o = cells + 12 * cid;
#pragma unroll
        for (int q = 0; q < 12; q++)
            o[q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    }
}

/* the cell-major copies of the context's grid(s), by its cells mode */
static int build_cells(tp_ctx *c)
{
    if (c->cells_mode == TP_CELLS_OFF) return TP_OK;
    const size_t nn = (size_t)c->f.nx * c->f.ny * c->f.nz;
    const int ns = c->grid_b ? 2 : 1;
    const size_t bytes = nn * 12 * sizeof(float4);
    if (c->cells_mode == TP_CELLS_AUTO) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess || ns * bytes > fr / 2) return TP_OK;
    }
    for (int s = 0; s < ns; s++) {
        float4 **dst = s == 0 ? &c->cells : &c->cells_b;
        if (hipMalloc((void **)dst, bytes) != hipSuccess) {
            *dst = nullptr;
            (void)hipFree(c->cells); c->cells = nullptr;
            return c->cells_mode == TP_CELLS_ON ? TP_E_NOMEM : TP_OK;
        }
        const size_t want = (nn + 255) / 256;
        const unsigned nb = (unsigned)(want < 65536 ? want : 65536);
        hipLaunchKernelGGL(k_build_cells, dim3(nb), dim3(256), 0, c->stream,
                           s == 0 ? c->grid : c->grid_b, c->f.nx, c->f.ny, c->f.nz, *dst);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->f.cells = c->cells;
    c->f.cells_b = c->cells_b;
    return TP_OK;
}

int tp_gpu_set_grid_cells(tp_ctx *c, int mode)
{
    if (!c || mode < TP_CELLS_AUTO || mode > TP_CELLS_ON) return TP_E_ARG;
    c->cells_mode = mode;
    return TP_OK;
}

namespace {
/* the context's two staging slabs for one load */
struct Staging {
    float *pin[2] = {nullptr, nullptr};   /* pinned host slabs */
    float *dev[2] = {nullptr, nullptr};   /* their device copies */
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
    size_t floats = 0;                    /* capacity of each, in floats */
    size_t rows = 0;                      /* slowest-axis rows per slab */
};

void staging_release(tp_ctx *c)
{
    for (int b = 0; b < 2; b++) {
        if (c->stage_pin[b]) (void)hipHostFree(c->stage_pin[b]);
        (void)hipFree(c->stage_dev[b]);
        if (c->stage_ev[b]) (void)hipEventDestroy(c->stage_ev[b]);
        c->stage_pin[b] = c->stage_dev[b] = nullptr;
        c->stage_ev[b] = nullptr;
    }
    c->stage_floats = 0;
}

/* slabs of at least `floats` floats: the context's, grown when too small */
int staging_get(tp_ctx *c, Staging &s, size_t floats)
{
    if (c->stage_floats < floats) {
        staging_release(c);
        for (int b = 0; b < 2; b++) {
            if (hipHostMalloc((void **)&c->stage_pin[b], floats * sizeof(float),
                              hipHostMallocDefault) != hipSuccess) {
                c->stage_pin[b] = nullptr;
                staging_release(c);
                return TP_E_NOMEM;
            }
            if (hipMalloc((void **)&c->stage_dev[b], floats * sizeof(float)) != hipSuccess) {
                c->stage_dev[b] = nullptr;
                staging_release(c);
                return TP_E_NOMEM;
            }
            if (hipEventCreateWithFlags(&c->stage_ev[b], hipEventDisableTiming) != hipSuccess) {
                c->stage_ev[b] = nullptr;
                staging_release(c);
                return TP_E_HIP;
            }
        }
        c->stage_floats = floats;
    }
    for (int b = 0; b < 2; b++) {
        s.pin[b] = c->stage_pin[b];
        s.dev[b] = c->stage_dev[b];
        s.ev[b] = c->stage_ev[b];
    }
    s.floats = c->stage_floats;
    return TP_OK;
}

/* One slice through the two staging slabs: while the GPU copies and lays
 * out slab b, the host fills slab b ^ 1 (after the event of its previous use).
 * The caller synchronizes the stream before the slabs are freed. */
int load_slice(tp_ctx *c, Staging &s, float4 *dst, int nx, int ny, int nz, int layout,
               double e0, double b0, tp_plane_reader rd, void *user, int slice)
{
    const size_t inner = layout == LAYOUT_XS ? (size_t)ny * nz : (size_t)nx * ny;
    const long long outer = layout == LAYOUT_XS ? nx : nz;
    const long long rows = (long long)s.rows;
    /* six planes of cnt floats, or two float3 arrays of 3*cnt */
    const int ncomp = layout == LAYOUT_AOS ? 2 : 6;
    int b = 0;
    for (long long s0 = 0; s0 < outer; s0 += rows, b ^= 1) {
        const long long s1 = s0 + rows < outer ? s0 + rows : outer;
        const size_t cnt = (size_t)(s1 - s0) * inner;
        const size_t per = cnt * (6 / ncomp);
        if (s.used[b]) HIPCHK(hipEventSynchronize(s.ev[b]));
        for (int comp = 0; comp < ncomp; comp++)
            if (rd(user, slice, comp, s0, s1, s.pin[b] + comp * per)) return TP_E_ARG;
        HIPCHK(hipMemcpyAsync(s.dev[b], s.pin[b], 6 * cnt * sizeof(float),
                              hipMemcpyHostToDevice, c->stream));
        const size_t want = (cnt + 255) / 256;
        const unsigned nb = (unsigned)(want < 65536 ? want : 65536);
        if (layout == LAYOUT_AOS) {
            hipLaunchKernelGGL(k_nodes_aos, dim3(nb), dim3(256), 0, c->stream, s.dev[b], cnt,
                               (size_t)s0 * inner, dst);
        } else if (layout == LAYOUT_XF) {
            hipLaunchKernelGGL(k_planes_xf, dim3(nb), dim3(256), 0, c->stream, s.dev[b], cnt,
                               (size_t)s0 * inner, dst, e0, b0);
        } else {
            const int ci = (int)(s1 - s0);
            dim3 g((nz + 31) / 32, (ci + 31) / 32, ny < 65535 ? ny : 65535);
            hipLaunchKernelGGL(k_planes_xs, g, dim3(256), 0, c->stream, s.dev[b], ci, (int)s0,
                               nx, ny, nz, dst, e0, b0);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(s.ev[b], c->stream));
        s.used[b] = true;
    }
    return TP_OK;
}

struct HostPlanes {
    const float *p[2];
    size_t nn, inner;
};

int host_planes_read(void *user, int slice, int comp, long long s0, long long s1, float *dst)
{
    const HostPlanes *h = (const HostPlanes *)user;
    memcpy(dst, h->p[slice] + comp * h->nn + (size_t)s0 * h->inner,
           (size_t)(s1 - s0) * h->inner * sizeof(float));
    return 0;
}

/* the float3 E and B arrays of one or two slices, read in z slabs */
struct HostNodes {
    const float *a[2][2];   /* [slice][E, B] */
    size_t inner;
};

int host_nodes_read(void *user, int slice, int comp, long long s0, long long s1, float *dst)
{
    const HostNodes *h = (const HostNodes *)user;
    memcpy(dst, h->a[slice][comp] + 3 * (size_t)s0 * h->inner,
           3 * (size_t)(s1 - s0) * h->inner * sizeof(float));
    return 0;
}

int load_grid(tp_ctx *c, int nx, int ny, int nz, const double lo[3], const double hi[3],
              int layout, int two_slices, double e0, double b0, double grid_dt, int mode,
              tp_plane_reader rd, void *user, size_t stage_bytes);
}  // namespace

int tp_gpu_load_field_grid(tp_ctx *c, int nx, int ny, int nz, const double lo[3],
                           const double hi[3], int order, int two_slices, double e0, double b0,
                           double grid_dt, int mode, tp_plane_reader rd, void *user,
                           size_t stage_bytes)
{
    if (order != TP_PLANES_X_FASTEST && order != TP_PLANES_X_SLOWEST) return TP_E_ARG;
    return load_grid(c, nx, ny, nz, lo, hi, order, two_slices, e0, b0, grid_dt, mode, rd, user,
                     stage_bytes);
}

int tp_gpu_set_field_grid(tp_ctx *c, int nx, int ny, int nz, const double lo[3],
                          const double hi[3], const float *ef, const float *bf, int mode)
{
    if (!ef || !bf) return TP_E_ARG;
    HostNodes h;
    h.a[0][0] = ef;
    h.a[0][1] = bf;
    h.a[1][0] = h.a[1][1] = nullptr;
    h.inner = (size_t)nx * ny;
    return load_grid(c, nx, ny, nz, lo, hi, LAYOUT_AOS, 0, 1.0, 1.0, 0.0, mode, host_nodes_read,
                     &h, 0);
}

int tp_gpu_set_field_mhd(tp_ctx *c, int nx, int ny, int nz, const double lo[3],
                         const double hi[3], const float *vf, const float *bf)
{
    /* the velocity takes the E slots of the nodes (and cell records); the
     * kernels form E = B x v after the gather */
    int rc = tp_gpu_set_field_grid(c, nx, ny, nz, lo, hi, vf, bf, TP_GATHER_REF_COMPAT);
    if (rc != TP_OK) return rc;
    c->f.kind = FK_MHD_REF;
    return TP_OK;
}

int tp_gpu_set_field_grid_double(tp_ctx *c, int nx, int ny, int nz, const double lo[3],
                                 const double hi[3], const double *ef, const double *bf)
{
    if (!c || !ef || !bf || !lo || !hi || nx < 2 || ny < 2 || nz < 2) return TP_E_ARG;
    const size_t nn = (size_t)nx * ny * nz;
    if (nn >= 0xFFFFFFFFull) return TP_E_ARG;    /* 32-bit corner indices */
    HIPCHK(hipSetDevice(c->device));
    free_field(c);
    memset(&c->f, 0, sizeof(c->f));
    if (hipMalloc(&c->gridd, nn * 8 * sizeof(double)) != hipSuccess) {
        c->gridd = nullptr;
        return TP_E_NOMEM;
    }
    /* interleaved on the host in slabs of at most 2^20 nodes */
    const size_t slab = nn < ((size_t)1 << 20) ? nn : ((size_t)1 << 20);
    double *h = (double *)malloc(slab * 8 * sizeof(double));
    if (!h) { free_field(c); return TP_E_NOMEM; }
    int rc = TP_OK;
    for (size_t n0 = 0; n0 < nn && rc == TP_OK; n0 += slab) {
        const size_t cnt = nn - n0 < slab ? nn - n0 : slab;
        for (size_t k = 0; k < cnt; k++) {
            const double *e = ef + 3 * (n0 + k), *b = bf + 3 * (n0 + k);
            double *o = h + 8 * k;
            o[0] = e[0]; o[1] = e[1]; o[2] = e[2];
            o[3] = b[0]; o[4] = b[1]; o[5] = b[2];
            o[6] = 0.0; o[7] = 0.0;
        }
        if (hipMemcpy(c->gridd + 8 * n0, h, cnt * 8 * sizeof(double), hipMemcpyHostToDevice) !=
            hipSuccess)
            rc = TP_E_HIP;
    }
    free(h);
    if (rc != TP_OK) { free_field(c); return rc; }
    c->f.nx = nx; c->f.ny = ny; c->f.nz = nz;
    set_box(c->f.lo, c->f.hi, c->f.d, lo, hi, nx, ny, nz);
    set_inv_d(c->f);
    c->f.gridd = c->gridd;
    c->f.kind = FK_GRIDD_REF;
    c->have_field = true;
    return TP_OK;
}

int tp_gpu_set_field_grid2(tp_ctx *c, int nx, int ny, int nz, const double lo[3],
                           const double hi[3], const float *ef_next, const float *bf_next,
                           const float *ef_prev, const float *bf_prev, double grid_dt,
                           int mode)
{
    if (!ef_next || !bf_next || !ef_prev || !bf_prev) return TP_E_ARG;
    HostNodes h;
    h.a[0][0] = ef_next;
    h.a[0][1] = bf_next;
    h.a[1][0] = ef_prev;
    h.a[1][1] = bf_prev;
    h.inner = (size_t)nx * ny;
    return load_grid(c, nx, ny, nz, lo, hi, LAYOUT_AOS, 1, 1.0, 1.0, grid_dt, mode,
                     host_nodes_read, &h, 0);
}

namespace {
int load_grid(tp_ctx *c, int nx, int ny, int nz, const double lo[3], const double hi[3],
              int layout, int two_slices, double e0, double b0, double grid_dt, int mode,
              tp_plane_reader rd, void *user, size_t stage_bytes)
{
    if (!c || nx < 2 || ny < 2 || nz < 2 || !lo || !hi || !rd) return TP_E_ARG;
    if (two_slices != 0 && two_slices != 1) return TP_E_ARG;
    if (two_slices && (!(grid_dt > 0.0) || !isfinite(grid_dt))) return TP_E_ARG;
    if (mode != TP_GATHER_REF_COMPAT && mode != TP_GATHER_TEXTBOOK) return TP_E_ARG;
    const size_t nn = (size_t)nx * ny * nz;
    if (nn >= 0xFFFFFFFFull) return TP_E_ARG;    /* 32-bit corner indices */
    HIPCHK(hipSetDevice(c->device));
    free_field(c);
    memset(&c->f, 0, sizeof(c->f));
    /* whole slabs of the six components, at least one, at most the grid */
    const size_t inner = layout == LAYOUT_XS ? (size_t)ny * nz : (size_t)nx * ny;
    const size_t outer = layout == LAYOUT_XS ? (size_t)nx : (size_t)nz;
    size_t rows = (stage_bytes ? stage_bytes : ((size_t)32 << 20)) / (6 * inner * sizeof(float));
    if (rows < 1) rows = 1;
    if (rows > outer) rows = outer;
    if (layout == LAYOUT_XS && rows > (size_t)65535 * 32) rows = (size_t)65535 * 32;
    if (hipMalloc(&c->grid, nn * 8 * sizeof(float)) != hipSuccess) {
        c->grid = nullptr;
        return TP_E_NOMEM;
    }
    if (two_slices && hipMalloc(&c->grid_b, nn * 8 * sizeof(float)) != hipSuccess) {
        c->grid_b = nullptr;
        free_field(c);
        return TP_E_NOMEM;
    }
    Staging s;
    int rc = staging_get(c, s, rows * 6 * inner);
    s.rows = rows;
    if (rc == TP_OK)
        rc = load_slice(c, s, c->grid, nx, ny, nz, layout, e0, b0, rd, user, 0);
    if (rc == TP_OK && two_slices)
        rc = load_slice(c, s, c->grid_b, nx, ny, nz, layout, e0, b0, rd, user, 1);
    /* every copy out of the slabs ends before the call returns (the next
     * load refills them) */
    const hipError_t se = hipStreamSynchronize(c->stream);
    if (rc == TP_OK && se != hipSuccess) rc = TP_E_HIP;
    if (rc != TP_OK) {
        free_field(c);
        memset(&c->f, 0, sizeof(c->f));
        return rc;
    }
    c->f.nx = nx; c->f.ny = ny; c->f.nz = nz;
    set_box(c->f.lo, c->f.hi, c->f.d, lo, hi, nx, ny, nz);
    set_inv_d(c->f);
    c->f.grid = c->grid;
    if (two_slices) {
        c->f.grid_b = c->grid_b;
        c->f.gdt = grid_dt;
        c->f.kind = mode == TP_GATHER_REF_COMPAT ? FK_GRID2_REF : FK_GRID2_TB;
    } else {
        c->f.kind = mode == TP_GATHER_REF_COMPAT ? FK_GRID_REF : FK_GRID_TB;
    }
    /* the reference gather reads cell records (the textbook gather's corner
     * order differs; it keeps reading nodes) */
    if (mode == TP_GATHER_REF_COMPAT) {
        rc = build_cells(c);
        if (rc != TP_OK) {
            free_field(c);
            memset(&c->f, 0, sizeof(c->f));
            return rc;
        }
    }
    c->have_field = true;
    return TP_OK;
}

}  // namespace

int tp_gpu_set_field_grid_planes(tp_ctx *c, int nx, int ny, int nz, const double lo[3],
                                 const double hi[3], int order, double e0, double b0,
                                 const float *planes_next, const float *planes_prev,
                                 double grid_dt, int mode)
{
    if (!planes_next || nx < 2 || ny < 2 || nz < 2) return TP_E_ARG;
    HostPlanes h;
    h.p[0] = planes_next;
    h.p[1] = planes_prev;
    h.nn = (size_t)nx * ny * nz;
    h.inner = order == TP_PLANES_X_FASTEST ? (size_t)nx * ny : (size_t)ny * nz;
    return tp_gpu_load_field_grid(c, nx, ny, nz, lo, hi, order, planes_prev ? 1 : 0, e0, b0,
                                  grid_dt, mode, host_planes_read, &h, 0);
}

int tp_gpu_get_field_nodes(tp_ctx *c, int slice, int nx, int ny, int nz, float *nodes)
{
    if (!c || !nodes || (slice != 0 && slice != 1)) return TP_E_ARG;
    const float4 *g = slice ? c->grid_b : c->grid;
    if (!c->have_field || !g) return TP_E_STATE;
    if (nx != c->f.nx || ny != c->f.ny || nz != c->f.nz) return TP_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    const size_t nn = (size_t)c->f.nx * c->f.ny * c->f.nz;
    HIPCHK(hipMemcpy(nodes, g, nn * 8 * sizeof(float), hipMemcpyDeviceToHost));
    return TP_OK;
}

int tp_gpu_set_field_wlcs(tp_ctx *c, int nwlcs, const double *rows, int ibessel,
                          double omega, double norm_i, int mode)
{
    if (!c || nwlcs < 1 || !rows) return TP_E_ARG;
    if (mode != TP_GATHER_REF_COMPAT && mode != TP_GATHER_TEXTBOOK) return TP_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    free_field(c);
    memset(&c->f, 0, sizeof(c->f));
    WlcsSys *h = (WlcsSys *)malloc(sizeof(WlcsSys) * nwlcs);
    if (!h) return TP_E_NOMEM;
    for (int i = 0; i < nwlcs; i++) {
        const double *r = rows + 14 * i;
        WlcsSys &s = h[i];
        s.cur_wr = r[0] * norm_i;
        s.x_wr = r[1]; s.y_wr = r[2]; s.z_wr = r[3];
        tp_sincos(r[4], &s.st, &s.ct);
        tp_sincos(r[5], &s.sp, &s.cp);
        s.t0 = r[6];
        s.cur_lp = r[7] * norm_i;
        s.r_lp = r[8]; s.x_lp = r[9]; s.y_lp = r[10]; s.z_lp = r[11];
        tp_sincos(r[12], &s.sa, &s.ca);
        tp_sincos(r[13], &s.sb, &s.cb);
        s.omega = omega;
    }
    hipError_t e = hipMalloc(&c->wl, sizeof(WlcsSys) * nwlcs);
    if (e != hipSuccess) { free(h); c->wl = nullptr; return TP_E_NOMEM; }
    e = hipMemcpy(c->wl, h, sizeof(WlcsSys) * nwlcs, hipMemcpyHostToDevice);
    free(h);
    if (e != hipSuccess) return TP_E_HIP;
    c->f.nwlcs = nwlcs;
    c->f.ibessel = ibessel;
    c->f.wlmode = mode == TP_GATHER_REF_COMPAT ? 0 : 1;
    c->f.wl = c->wl;
    c->f.kind = FK_WLCS;
    c->have_field = true;
    return TP_OK;
}

int tp_gpu_set_field_wlcs_table(tp_ctx *c, int nwlcs, const double *tab, int ibessel,
                                int mode)
{
    if (!c || nwlcs < 1 || !tab) return TP_E_ARG;
    if (mode != TP_GATHER_REF_COMPAT && mode != TP_GATHER_TEXTBOOK) return TP_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    free_field(c);
    memset(&c->f, 0, sizeof(c->f));
    WlcsSys *h = (WlcsSys *)malloc(sizeof(WlcsSys) * nwlcs);
    if (!h) return TP_E_NOMEM;
    for (int i = 0; i < nwlcs; i++) {
        const double *r = tab + 19 * i;
        WlcsSys &s = h[i];
        s.x_wr = r[0]; s.y_wr = r[1]; s.z_wr = r[2];
        s.ct = r[3]; s.cp = r[4]; s.st = r[5]; s.sp = r[6];
        s.cur_wr = r[7]; s.t0 = r[8];
        s.x_lp = r[9]; s.y_lp = r[10]; s.z_lp = r[11]; s.r_lp = r[12];
        s.ca = r[13]; s.cb = r[14]; s.sa = r[15]; s.sb = r[16];
        s.cur_lp = r[17]; s.omega = r[18];
    }
    hipError_t e = hipMalloc(&c->wl, sizeof(WlcsSys) * nwlcs);
    if (e != hipSuccess) { free(h); c->wl = nullptr; return TP_E_NOMEM; }
    e = hipMemcpy(c->wl, h, sizeof(WlcsSys) * nwlcs, hipMemcpyHostToDevice);
    free(h);
    if (e != hipSuccess) return TP_E_HIP;
    c->f.nwlcs = nwlcs;
    c->f.ibessel = ibessel;
    c->f.wlmode = mode == TP_GATHER_REF_COMPAT ? 0 : 1;
    c->f.wl = c->wl;
    c->f.kind = FK_WLCS;
    c->have_field = true;
    return TP_OK;
}

int tp_gpu_set_field_modes(tp_ctx *c, int nmodes, const double *rows, double b0z)
{
    if (!c || nmodes < 0 || (nmodes > 0 && !rows)) return TP_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    free_field(c);
    memset(&c->f, 0, sizeof(c->f));
    if (nmodes > 0) {
        size_t bytes = sizeof(double) * 7 * nmodes;
        if (hipMalloc(&c->modes, bytes) != hipSuccess) { c->modes = nullptr; return TP_E_NOMEM; }
        HIPCHK(hipMemcpy(c->modes, rows, bytes, hipMemcpyHostToDevice));
        /* FP32 copy: (kx ky kz phi) / 2 pi, ax ay az 0 */
        float *t32 = (float *)malloc(sizeof(float) * 8 * nmodes);
        if (!t32) return TP_E_NOMEM;
        const double inv2pi = 1.0 / (2.0 * TP_PI);
        for (int m = 0; m < nmodes; m++) {
            const double *r = rows + 7 * m;
            float *o = t32 + 8 * m;
            o[0] = (float)(r[0] * inv2pi); o[1] = (float)(r[1] * inv2pi);
            o[2] = (float)(r[2] * inv2pi); o[3] = (float)(r[6] * inv2pi);
            o[4] = (float)r[3]; o[5] = (float)r[4]; o[6] = (float)r[5]; o[7] = 0.f;
        }
        if (hipMalloc(&c->modes32, sizeof(float) * 8 * nmodes) != hipSuccess) {
            c->modes32 = nullptr;
            free(t32);
            return TP_E_NOMEM;
        }
        hipError_t e = hipMemcpy(c->modes32, t32, sizeof(float) * 8 * nmodes,
                                 hipMemcpyHostToDevice);
        free(t32);
        HIPCHK(e);
    }
    c->f.nmodes = nmodes;
    c->f.modes = c->modes;
    c->f.modes32 = c->modes32;
    c->f.b0z = b0z;
    c->f.kind = FK_MODES;
    c->have_field = true;
    return TP_OK;
}

int tp_gpu_eval_field(tp_ctx *c, int n, const double *pts, double *out, int prec)
{
    if (!c || n < 0 || (n > 0 && (!pts || !out))) return TP_E_ARG;
    if (!c->have_field) return TP_E_STATE;
    if (n == 0) return TP_OK;
    HIPCHK(hipSetDevice(c->device));
    double *dp = nullptr, *dout = nullptr;
    if (hipMalloc(&dp, sizeof(double) * 4 * n) != hipSuccess) return TP_E_NOMEM;
    if (hipMalloc(&dout, sizeof(double) * 6 * n) != hipSuccess) { (void)hipFree(dp); return TP_E_NOMEM; }
    int rc = TP_OK;
    if (hipMemcpy(dp, pts, sizeof(double) * 4 * n, hipMemcpyHostToDevice) != hipSuccess) rc = TP_E_HIP;
    if (rc == TP_OK) {
        dispatch_eval(field_kind(c), dim3((n + 255) / 256), dim3(256), c->stream, c->f, n, dp,
                      dout, prec);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
            rc = TP_E_HIP;
    }
    if (rc == TP_OK &&
        hipMemcpy(out, dout, sizeof(double) * 6 * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = TP_E_HIP;
    (void)hipFree(dp);
    (void)hipFree(dout);
    return rc;
}

static int ensure_state(tp_ctx *c, int n)
{
    State &S = c->s;
    if (n <= S.cap) { S.n = n; return TP_OK; }
    (void)hipFree(S.buf); (void)hipFree(S.ibuf); (void)hipFree(S.stage);
    S.buf = nullptr; S.ibuf = nullptr; S.stage = nullptr; S.cap = 0;
    size_t cap = (size_t)n;
    if (hipMalloc(&S.buf, sizeof(double) * 25 * cap) != hipSuccess) { S.buf = nullptr; return TP_E_NOMEM; }
    if (hipMalloc(&S.ibuf, sizeof(int) * 2 * cap) != hipSuccess) { S.ibuf = nullptr; return TP_E_NOMEM; }
    if (hipMalloc(&S.stage, sizeof(tp_particle) * 2 * cap) != hipSuccess) { S.stage = nullptr; return TP_E_NOMEM; }
    double *b = S.buf;
    double **arr[20] = {&S.x, &S.y, &S.z, &S.ux, &S.uy, &S.uz, &S.t, &S.xs, &S.ys, &S.zs,
                        &S.x0, &S.y0, &S.z0, &S.ux0, &S.uy0, &S.uz0, &S.t0, &S.g0, &S.mu0, &S.th0};
    for (int k = 0; k < 20; k++) *arr[k] = b + (size_t)k * cap;
    S.st = b + 20 * cap;   /* 5 * cap */
    S.esc = S.ibuf;
    S.nsteps = S.ibuf + cap;
    S.cap = n;
    S.n = n;
    return TP_OK;
}

int tp_gpu_load_particles(tp_ctx *c, int n, const tp_particle *ptl, const tp_particle *ptl_init)
{
    if (!c || n < 0 || (n > 0 && (!ptl || !ptl_init))) return TP_E_ARG;
    if (!c->have_field) return TP_E_STATE;
    HIPCHK(hipSetDevice(c->device));
    int rc = ensure_state(c, n);
    if (rc) return rc;
    c->s.derived = false;
    if (n == 0) {
        c->s.derived = true;
        return TP_OK;
    }
    HIPCHK(hipMemcpyAsync(c->s.stage, ptl, sizeof(tp_particle) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->s.stage + n, ptl_init, sizeof(tp_particle) * n, hipMemcpyHostToDevice,
                          c->stream));
    dispatch_load(field_kind(c), dim3((n + 255) / 256), dim3(256), c->stream, c->f, n,
                  c->s.stage, c->s.stage + n, c->s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->s.derived = true;
    return TP_OK;
}

static int block_size_for(const tp_ctx *c, int n)
{
    /* 256 lanes per block unless that leaves compute units idle */
    int cu = c->num_cu > 0 ? c->num_cu : 256;
    if ((n + 255) / 256 >= 2 * cu) return 256;
    if ((n + 127) / 128 >= 2 * cu) return 128;
    return 64;
}

/* Buffers of one fixed-step run: the context's own (tp_gpu_run_fixed) or the
 * caller's device outputs and the static scratch array
 * (tp_gpu_run_fixed_noalloc). */
struct RunBufs {
    unsigned *espect, *espect_esc;  /* [nt_out][nbins] counts */
    double *dco;                    /* [8][nd] D sums */
    int *rdec;                      /* [nd] escapes counted against each D row */
    double *partials;               /* block rows of 8 partial D sums */
    size_t prows;
    int *live;                      /* open boundaries: the index lists */
    size_t live_ints;
    int *flag;                      /* non-finite state found (1 int) */
    bool per_launch_events;         /* c->evs around every launch, else ev0/ev1
                                       around the run */
};

__global__ void __launch_bounds__(256) k_fill_u32(unsigned *a, size_t n, unsigned v)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        a[i] = v;
}

__global__ void __launch_bounds__(256) k_fill_f64(double *a, size_t n, double v)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        a[i] = v;
}

/* the outputs a fixed-step run accumulates into, cleared in one launch
 * (one dispatch instead of four ahead of the step kernel): counts
 * [0, na) and [na, 2 na), D sums [2 na, 2 na + 8 nd), escape counts per D
 * row after that */
__global__ void __launch_bounds__(256) k_clear_run(unsigned *espect, unsigned *espect_esc,
                                                   size_t na, double *dco, int *rdec, size_t nd)
{
    const size_t total = 2 * na + 9 * nd;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total;
         i += (size_t)gridDim.x * 256) {
        if (i < na) espect[i] = 0u;
        else if (i < 2 * na) espect_esc[i - na] = 0u;
        else if (i < 2 * na + 8 * nd) dco[i - 2 * na] = 0.0;
        else rdec[i - 2 * na - 8 * nd] = 0;
    }
}

/* remaining counts from the escapes counted against each D row, in place */
__global__ void __launch_bounds__(256) k_remain(int *rdec, int nd, int n)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < nd) rdec[k] = n - rdec[k];
}

/* flag = 1 if any particle's position or momentum is NaN or infinite (the
 * reference lets them propagate: B = 0 makes iBtot infinite in
 * tracking_wirz, src/tracking.c:782) */
__global__ void __launch_bounds__(256) k_nonfinite(int n, const double *x, const double *y,
                                                   const double *z, const double *ux,
                                                   const double *uy, const double *uz,
                                                   int *flag)
{
    bool bad = false;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        bad |= !isfinite(x[i]) || !isfinite(y[i]) || !isfinite(z[i]) || !isfinite(ux[i]) ||
               !isfinite(uy[i]) || !isfinite(uz[i]);
    if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

static int fill_u32(hipStream_t s, unsigned *a, size_t n)
{
    if (n == 0) return TP_OK;
    const size_t nbk = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)nbk), dim3(256), 0, s, a, n, 0u);
    return hipGetLastError() == hipSuccess ? TP_OK : TP_E_HIP;
}

static int fill_f64(hipStream_t s, double *a, size_t n)
{
    if (n == 0) return TP_OK;
    const size_t nbk = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)nbk), dim3(256), 0, s, a, n, 0.0);
    return hipGetLastError() == hipSuccess ? TP_OK : TP_E_HIP;
}

/* the checks of a run's parameters; the step kernel that runs it */
static int run_check(const tp_ctx *c, const tp_track_params *p, track_fn *fn, bool *mlp)
{
    if (!c || !p) return TP_E_ARG;
    if (!c->have_field || !c->s.derived) return TP_E_STATE;
    if (p->einterval_t <= 0 || p->dcoeff_interval <= 0 || p->estep < p->sstep ||
        p->nbins < 1 || p->nt_out < 1 || p->sstep < 0 || !(p->dt > 0.0) ||
        p->bc_flag < 0 || p->bc_flag > 2)
        return TP_E_ARG;
    if (p->pusher != TP_PUSH_WIRZ && p->pusher != TP_PUSH_BORIS) return TP_E_ARG;
    if (p->precision != TP_FP64 && p->precision != TP_FP32_MIXED) return TP_E_ARG;
    *mlp = field_kind(c) == FK_MLP;
    if (*mlp && (p->pusher != TP_PUSH_BORIS || p->precision != TP_FP32_MIXED)) return TP_E_ARG;
    if (*mlp && c->traj_nout > 0) return TP_E_ARG;
    /* surrogate field: the exact-field kernel runs the particles that fell back */
    *fn = pick_track(*mlp ? FK_MODES : field_kind(c), p->pusher, p->precision,
                     c->traj_nout > 0);
    return *fn ? TP_OK : TP_E_ARG;
}

/* the sizes a run over n particles needs */
struct RunSizes {
    int nd, bs, nblocks, nb256;
    size_t prows, live_ints;
    long nlaunch;
};

static RunSizes run_sizes(const tp_ctx *c, const tp_track_params *p, int n)
{
    RunSizes r;
    const int di = p->dcoeff_interval;
    r.nd = p->estep / di + 1;
    r.bs = block_size_for(c, n);
    r.nblocks = n > 0 ? (n + r.bs - 1) / r.bs : 0;
    r.nb256 = (n + 255) / 256;
    /* partial-sum rows: the step kernel's blocks, or with open boundaries a
     * launch over the particles inside plus k_dstale's blocks of 256 */
    r.prows = (size_t)(r.nblocks > 0 ? r.nblocks : 1) + r.nb256 + 2;
    /* open boundaries: inside[n], gone[n], block counts and offsets, total */
    r.live_ints = 2 * (size_t)n + 2 * (size_t)r.nb256 + 1;
    /* launches: each ends at the next D step (a multiple of di) or estep */
    r.nlaunch = 0;
    for (long j0 = p->sstep; j0 <= p->estep; r.nlaunch++) {
        long j1 = (j0 / di + 1) * (long)di;
        if (j0 % di == 0) j1 = j0;
        if (j1 > p->estep) j1 = p->estep;
        j0 = j1 + 1;
    }
    return r;
}

/* the launches of a fixed-step run over the resident particles into the
 * buffers B (cleared here) */
static int run_fixed_impl(tp_ctx *c, const tp_track_params *p, track_fn fn, bool mlp,
                          const RunSizes &R, const RunBufs &B)
{
    const int n = c->s.n;
    const int nb = p->nbins, nt = p->nt_out, di = p->dcoeff_interval;
    const int nd = R.nd, bs = R.bs, nblocks = R.nblocks, nb256 = R.nb256;
    const size_t need = (size_t)nb * nt;
    int rc;
    {
        const size_t total = 2 * need + 9 * (size_t)nd;
        const size_t nbk = total ? (total + 255) / 256 < 1024 ? (total + 255) / 256 : 1024 : 1;
        hipLaunchKernelGGL(k_clear_run, dim3((unsigned)nbk), dim3(256), 0, c->stream, B.espect,
                           B.espect_esc, need, B.dco, B.rdec, (size_t)nd);
        HIPCHK(hipGetLastError());
    }
    c->last_ms = 0.0;
    c->last_launches = 0;

    size_t shmem = sizeof(unsigned) * nb;
    if (shmem > 64 * 1024) return TP_E_ARG;
    /* open boundaries: after each launch the particles are split (stably)
     * into those still inside and those that left, and once some have left
     * the next launch runs over the first list, so the lanes of escaped
     * particles do not idle through the remaining steps.  Its D terms are
     * formed in the step kernel for the particles inside and by k_dstale
     * (stale terms, escape count) for the others; per-particle results and
     * spectra are unchanged, the D sums are summed in that block order. */
    const bool compact = p->bc_flag == 1 && !mlp && n > 0;
    int *l_in = B.live, *l_gone = l_in + n, *l_cnt = l_gone + n;
    int *l_off = l_cnt + nb256, *l_tot = l_off + nb256;

    TrackArgs a;
    memset(&a, 0, sizeof(a));
    a.f = c->f;
    /* periodic particles in the grid's own box stay in it between steps */
    a.f.in_box = p->bc_flag == 0 && c->f.lo[0] == p->xmin && c->f.hi[0] == p->xmax &&
                 c->f.lo[1] == p->ymin && c->f.hi[1] == p->ymax && c->f.lo[2] == p->zmin &&
                 c->f.hi[2] == p->zmax;
    a.n = n; a.ntot = n; a.pboff = 0; a.bc = p->bc_flag; a.nbins = nb; a.nt_out = nt; a.einterval = p->einterval_t;
    a.di = di; a.fix_stale = (p->flags & TP_FLAG_FIX_STALE_DCOEFF) ? 1 : 0;
    a.nd = nd;
    a.dt = p->dt; a.qm = p->qm;
    a.lo[0] = p->xmin; a.lo[1] = p->ymin; a.lo[2] = p->zmin;
    a.hi[0] = p->xmax; a.hi[1] = p->ymax; a.hi[2] = p->zmax;
    a.logemin = p->logemin; a.logde = p->logde;
    State &S = c->s;
    a.x = S.x; a.y = S.y; a.z = S.z; a.ux = S.ux; a.uy = S.uy; a.uz = S.uz; a.t = S.t;
    a.xs = S.xs; a.ys = S.ys; a.zs = S.zs;
    a.x0 = S.x0; a.y0 = S.y0; a.z0 = S.z0; a.ux0 = S.ux0; a.uy0 = S.uy0; a.uz0 = S.uz0;
    a.g0 = S.g0; a.mu0 = S.mu0; a.th0 = S.th0;
    a.st = S.st; a.esc = S.esc; a.nsteps = S.nsteps;
    a.espect = B.espect; a.espect_esc = B.espect_esc;
    a.partials = B.partials; a.rdec = B.rdec;
    if (c->traj_nout > 0) {
        a.traj = c->traj; a.traj_accum = c->traj_accum; a.traj_pts = c->traj_pts;
        a.traj_nout = c->traj_nout; a.stage = S.stage;
    }

    /* frame and D-row counters at step sstep (src/tracking.c:317, 328) */
    int ntp = (int)floor(p->sstep / (double)p->einterval_t);
    int tstep = 1;
    int launches = 0;
    bool use_list = false;      /* the launch runs over l_in[0 .. nlive) */
    int nlive = n;
    int row_blocks = nblocks;   /* partial-sum rows of the launch's D row */
    if (!B.per_launch_events) HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (long j0 = p->sstep; j0 <= p->estep && nblocks > 0;) {
        long j1 = (j0 % di == 0) ? j0 : (j0 / di + 1) * (long)di;
        if (j1 > p->estep) j1 = p->estep;
        a.j0 = (int)j0; a.j1 = (int)j1;
        a.first = (j0 == p->sstep) ? 1 : 0;
        a.ntp0 = ntp;
        /* the D terms in the step kernel, unless it runs over index lists */
        const bool dstep = j1 % di == 0;
        a.dstep = (dstep && !mlp) ? tstep : -1;
        row_blocks = nblocks;
        if (B.per_launch_events) HIPCHK(hipEventRecord(c->evs[2 * launches], c->stream));
        if (use_list) {
            const int nbl = (nlive + bs - 1) / bs, nesc = n - nlive;
            if (nlive > 0) {
                TrackArgs b = a;
                b.idx = l_in;
                b.n = nlive;
                fn(dim3(nbl), dim3(bs), shmem, c->stream, b);
                HIPCHK(hipGetLastError());
            }
            row_blocks = nbl;
            if (dstep && nesc > 0) {
                TrackArgs b = a;
                b.idx = l_gone;
                b.n = nesc;
                b.pboff = nbl;
                hipLaunchKernelGGL(k_dstale, dim3((nesc + 255) / 256), dim3(256), 0,
                                   c->stream, b, tstep);
                row_blocks = nbl + (nesc + 255) / 256;
            }
        } else if (mlp) {
            /* residual checks before the first launch and at launches that
             * start after a multiple of the check interval */
            const int ci = c->f.mlp.check;
            const int chk = a.first || (ci > 0 && (j0 - 1) % ci == 0);
            rc = tp_mlp_launch(c, a, shmem, chk, bs, fn);
            if (rc) return rc;
        } else {
            fn(dim3(nblocks), dim3(bs), shmem, c->stream, a);
        }
        HIPCHK(hipGetLastError());
        if (B.per_launch_events) HIPCHK(hipEventRecord(c->evs[2 * launches + 1], c->stream));
        launches++;
        for (long j = j0; j <= j1; j++)
            if (j % p->einterval_t == 0) ntp++;
        if (dstep) {
            if (a.dstep < 0) {
                dispatch_dcoeff(field_kind(c), dim3(nblocks), dim3(bs), c->stream, a, tstep);
                HIPCHK(hipGetLastError());
            }
            hipLaunchKernelGGL(k_reduce_row, dim3(8), dim3(256), 0, c->stream, B.partials,
                               row_blocks, tstep, nd, B.dco);
            HIPCHK(hipGetLastError());
            tstep++;
        }
        if (compact && j1 < p->estep) {
            hipLaunchKernelGGL(k_live_count, dim3(nb256), dim3(256), 0, c->stream,
                               (const int *)S.esc, n, l_cnt);
            hipLaunchKernelGGL(k_live_scan, dim3(1), dim3(1024), 0, c->stream,
                               (const int *)l_cnt, nb256, l_off, l_tot);
            hipLaunchKernelGGL(k_live_scatter, dim3(nb256), dim3(256), 0, c->stream,
                               (const int *)S.esc, n, (const int *)l_off, l_in, l_gone);
            HIPCHK(hipGetLastError());
            int h = 0;
            HIPCHK(hipMemcpyAsync(&h, l_tot, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            if (h < 0 || h > n) return TP_E_HIP;
            nlive = h;
            use_list = h < n;
        }
        j0 = j1 + 1;
    }
    if (!B.per_launch_events) HIPCHK(hipEventRecord(c->ev1, c->stream));
    int bad = 0;
    if ((p->flags & TP_FLAG_CHECK_FINITE) && n > 0) {
        if ((rc = fill_u32(c->stream, (unsigned *)B.flag, 1))) return rc;
        const int nbk = nb256 < 2048 ? nb256 : 2048;
        hipLaunchKernelGGL(k_nonfinite, dim3(nbk), dim3(256), 0, c->stream, n,
                           (const double *)S.x, (const double *)S.y, (const double *)S.z,
                           (const double *)S.ux, (const double *)S.uy, (const double *)S.uz,
                           B.flag);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(&bad, B.flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    double ms_total = 0.0;
    if (B.per_launch_events) {
        for (int k = 0; k < launches; k++) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->evs[2 * k], c->evs[2 * k + 1]));
            ms_total += ms;
        }
    } else if (launches > 0) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        ms_total = ms;
    }
    c->last_ms = ms_total;
    c->last_launches = launches;
    return bad ? TP_E_NONFINITE : TP_OK;
}

int tp_gpu_run_fixed(tp_ctx *c, const tp_track_params *p)
{
    track_fn fn = nullptr;
    bool mlp = false;
    int rc = run_check(c, p, &fn, &mlp);
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    const int n = c->s.n;
    if (mlp) {
        rc = tp_mlp_prepare_run(c, n);
        if (rc) return rc;
    }
    const RunSizes R = run_sizes(c, p, n);
    const int nd = R.nd;
    /* diagnostics buffers */
    size_t need = (size_t)p->nbins * p->nt_out;
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
    /* partial-sum rows and the non-finite flag after them */
    size_t pneed = R.prows * 8 + 1;
    if (pneed > c->partials_cap) {
        (void)hipFree(c->partials);
        c->partials = nullptr; c->partials_cap = 0;
        if (hipMalloc(&c->partials, pneed * sizeof(double)) != hipSuccess) return TP_E_NOMEM;
        c->partials_cap = pneed;
    }
    const bool compact = p->bc_flag == 1 && !mlp && n > 0;
    if (compact && R.live_ints > c->live_cap) {
        (void)hipFree(c->live_idx);
        c->live_idx = nullptr; c->live_cap = 0;
        if (hipMalloc(&c->live_idx, R.live_ints * sizeof(int)) != hipSuccess) return TP_E_NOMEM;
        c->live_cap = R.live_ints;
    }
    if (2 * R.nlaunch > c->nev) {      /* two events per launch */
        for (int k = 0; k < c->nev; k++) (void)hipEventDestroy(c->evs[k]);
        free(c->evs);
        c->nev = 0;
        c->evs = (hipEvent_t *)malloc(sizeof(hipEvent_t) * 2 * R.nlaunch);
        if (!c->evs) return TP_E_NOMEM;
        for (long k = 0; k < 2 * R.nlaunch; k++) {
            HIPCHK(hipEventCreate(&c->evs[k]));
            c->nev = (int)(k + 1);
        }
        c->nev = (int)(2 * R.nlaunch);
    }
    c->nb = p->nbins; c->nt = p->nt_out; c->nd = nd; c->last_n = n;
    RunBufs B;
    B.espect = c->espect; B.espect_esc = c->espect_esc;
    B.dco = c->dco; B.rdec = c->rdec;
    B.partials = c->partials; B.prows = R.prows;
    B.live = c->live_idx; B.live_ints = c->live_cap;
    B.flag = (int *)(c->partials + R.prows * 8);
    B.per_launch_events = true;
    return run_fixed_impl(c, p, fn, mlp, R, B);
}

/* ------------------------------------------------------------------ */
/* the allocation-free run: scratch in a static device array           */
/* ------------------------------------------------------------------ */
/* Scratch of tp_gpu_run_fixed_noalloc: the D partial-sum rows, the
 * open-boundary index lists and the non-finite flag of runs over at most
 * TP_NOALLOC_MAX_PARTICLES particles.  Every GPU the library runs on holds
 * one copy (about 36 MB) until the process ends; host code reaches it only
 * through the device address kept per device by noalloc_scratch(). */
#define NOALLOC_BLK_ROWS (2 * ((TP_NOALLOC_MAX_PARTICLES + 255) / 256) + 2)
#define NOALLOC_LIVE_INTS (2 * (size_t)TP_NOALLOC_MAX_PARTICLES + \
                           2 * ((TP_NOALLOC_MAX_PARTICLES + 255) / 256) + 1)
#define NOALLOC_DOUBLES ((size_t)NOALLOC_BLK_ROWS * 8 + (NOALLOC_LIVE_INTS + 1) / 2 + 1)
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

#define NOALLOC_MAX_DEV 64
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[NOALLOC_MAX_DEV];
static int g_scratch_rc[NOALLOC_MAX_DEV];   /* 0 not looked up, 1 ok, -1 failed */

/* The device address of scratch on the current device `dev`, looked up once
 * per device (each GPU holds its own copy at its own address). */
static void *noalloc_scratch(int dev)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEV) return nullptr;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_rc[dev] == 0) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            fprintf(stderr, "tp_gpu: hipGetSymbolAddress(scratch) on device %d failed: %s\n",
                    dev, hipGetErrorString(e));
            g_scratch_rc[dev] = -1;
        } else {
            g_scratch_ptr[dev] = ptr;
            g_scratch_rc[dev] = 1;
        }
    }
    return g_scratch_rc[dev] == 1 ? g_scratch_ptr[dev] : nullptr;
}

int tp_gpu_run_fixed_noalloc(tp_ctx *c, const tp_track_params *p, unsigned *espect,
                             unsigned *espect_escape, double *dcoeff, int *nptl_remain)
{
    track_fn fn = nullptr;
    bool mlp = false;
    int rc = run_check(c, p, &fn, &mlp);
    if (rc) return rc;
    if (mlp || c->traj_nout > 0) return TP_E_ARG;   /* these paths allocate */
    if (!espect || !espect_escape || !dcoeff || !nptl_remain) return TP_E_ARG;
    const int n = c->s.n;
    if (n > TP_NOALLOC_MAX_PARTICLES) {
        fprintf(stderr, "tp_gpu_run_fixed_noalloc: %d particles, at most %d\n", n,
                TP_NOALLOC_MAX_PARTICLES);
        return TP_E_ARG;
    }
    const RunSizes R = run_sizes(c, p, n);
    if (R.prows > (size_t)NOALLOC_BLK_ROWS || R.live_ints > NOALLOC_LIVE_INTS) {
        fprintf(stderr, "tp_gpu_run_fixed_noalloc: the run needs more scratch than the "
                        "library holds\n");
        return TP_E_ARG;
    }
    HIPCHK(hipSetDevice(c->device));
    double *sp = c->scratch;
    if (!sp) return TP_E_HIP;
    RunBufs B;
    B.espect = espect; B.espect_esc = espect_escape;
    B.dco = dcoeff; B.rdec = nptl_remain;
    B.partials = sp; B.prows = R.prows;
    B.live = (int *)(sp + (size_t)NOALLOC_BLK_ROWS * 8);
    B.live_ints = NOALLOC_LIVE_INTS;
    B.flag = (int *)(sp + NOALLOC_DOUBLES - 1);
    B.per_launch_events = false;
    rc = run_fixed_impl(c, p, fn, false, R, B);
    if (rc != TP_OK && rc != TP_E_NONFINITE) return rc;
    /* the escapes counted against each row become the remaining counts */
    hipLaunchKernelGGL(k_remain, dim3((R.nd + 255) / 256), dim3(256), 0, c->stream,
                       nptl_remain, R.nd, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return rc;
}

int tp_gpu_last_kernel_time(tp_ctx *c, double *ms, int *launches)
{
    if (!c) return TP_E_ARG;
    if (ms) *ms = c->last_ms;
    if (launches) *launches = c->last_launches;
    return TP_OK;
}

int tp_gpu_store_particles(tp_ctx *c, tp_particle *ptl, int *nsteps)
{
    if (!c) return TP_E_ARG;
    const int n = c->s.n;
    if (n == 0) return TP_OK;
    if (!c->s.buf) return TP_E_STATE;
    HIPCHK(hipSetDevice(c->device));
    if (ptl) {
        State &S = c->s;
        hipLaunchKernelGGL(k_store, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, S.x, S.y,
                           S.z, S.ux, S.uy, S.uz, S.t, S.xs, S.ys, S.zs, S.stage);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(ptl, S.stage, sizeof(tp_particle) * n, hipMemcpyDeviceToHost,
                              c->stream));
    }
    if (nsteps)
        HIPCHK(hipMemcpyAsync(nsteps, c->s.nsteps, sizeof(int) * n, hipMemcpyDeviceToHost,
                              c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return TP_OK;
}

int tp_gpu_read_diagnostics(tp_ctx *c, double *espect, double *espect_escape, double *dcoeff,
                            int *nptl_remain)
{
    if (!c) return TP_E_ARG;
    if (!c->espect || !c->dco) return TP_E_STATE;
    HIPCHK(hipSetDevice(c->device));
    const size_t nbt = (size_t)c->nb * c->nt;
    unsigned *h = (unsigned *)malloc(sizeof(unsigned) * (nbt > 0 ? nbt : 1));
    int *r = (int *)malloc(sizeof(int) * (c->nd + 1));
    if (!h || !r) { free(h); free(r); return TP_E_NOMEM; }
    int rc = TP_OK;
    if (espect) {
        if (hipMemcpy(h, c->espect, sizeof(unsigned) * nbt, hipMemcpyDeviceToHost) != hipSuccess)
            rc = TP_E_HIP;
        for (size_t k = 0; rc == TP_OK && k < nbt; k++) espect[k] = (double)h[k];
    }
    if (rc == TP_OK && espect_escape) {
        if (hipMemcpy(h, c->espect_esc, sizeof(unsigned) * nbt, hipMemcpyDeviceToHost) != hipSuccess)
            rc = TP_E_HIP;
        for (size_t k = 0; rc == TP_OK && k < nbt; k++) espect_escape[k] = (double)h[k];
    }
    if (rc == TP_OK && dcoeff &&
        hipMemcpy(dcoeff, c->dco, sizeof(double) * 8 * c->nd, hipMemcpyDeviceToHost) != hipSuccess)
        rc = TP_E_HIP;
    if (rc == TP_OK && nptl_remain) {
        if (hipMemcpy(r, c->rdec, sizeof(int) * c->nd, hipMemcpyDeviceToHost) != hipSuccess)
            rc = TP_E_HIP;
        for (int k = 0; rc == TP_OK && k < c->nd; k++) nptl_remain[k] = c->last_n - r[k];
    }
    free(h);
    free(r);
    return rc;
}

__global__ void k_fill_int(int *a, int n, int v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

/* trajectory buffers of the next run: the caller's slots (ptl_time, slot
 * ranges ntraj_accum) on the device, every particle's next slot 1 */
static int traj_arm(tp_ctx *c, int n, int nsteps_output, const int *ntraj_accum,
                    const tp_particle *ptl_time)
{
    if (n < 0 || nsteps_output <= 0 || (n > 0 && (!ntraj_accum || !ptl_time))) return TP_E_ARG;
    if (field_kind(c) == FK_MLP) return TP_E_ARG;
    /* slot ranges must be non-decreasing from 0 */
    for (int i = 0; i < n; i++)
        if (ntraj_accum[i] < (i > 0 ? ntraj_accum[i - 1] : 0)) return TP_E_ARG;
    const size_t ntraj = n > 0 ? (size_t)ntraj_accum[n - 1] : 0;
    HIPCHK(hipSetDevice(c->device));
    if (ntraj > c->traj_cap || n > c->traj_ncap) {
        (void)hipFree(c->traj); (void)hipFree(c->traj_accum); (void)hipFree(c->traj_pts);
        c->traj = nullptr; c->traj_accum = nullptr; c->traj_pts = nullptr;
        c->traj_cap = 0; c->traj_ncap = 0;
        const size_t t = ntraj > 0 ? ntraj : 1, k = n > 0 ? (size_t)n : 1;
        if (hipMalloc(&c->traj, t * sizeof(tp_particle)) != hipSuccess ||
            hipMalloc(&c->traj_accum, k * sizeof(int)) != hipSuccess ||
            hipMalloc(&c->traj_pts, k * sizeof(int)) != hipSuccess)
            return TP_E_NOMEM;
        c->traj_cap = t;
        c->traj_ncap = (int)k;
    }
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(c->traj, ptl_time, ntraj * sizeof(tp_particle),
                              hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->traj_accum, ntraj_accum, n * sizeof(int),
                              hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                           c->traj_pts, n, 1);
        HIPCHK(hipGetLastError());
    }
    c->traj_nout = nsteps_output;
    return TP_OK;
}

/* the armed run's records back to the caller (and disarm) */
static int traj_collect(tp_ctx *c, int n, const int *ntraj_accum, tp_particle *ptl_time,
                        int *points)
{
    c->traj_nout = 0;
    if (n > 0) {
        const size_t ntraj = (size_t)ntraj_accum[n - 1];
        HIPCHK(hipMemcpyAsync(ptl_time, c->traj, ntraj * sizeof(tp_particle),
                              hipMemcpyDeviceToHost, c->stream));
        if (points)
            HIPCHK(hipMemcpyAsync(points, c->traj_pts, n * sizeof(int), hipMemcpyDeviceToHost,
                                  c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return TP_OK;
}

int tp_gpu_track_fixed_traj(tp_ctx *c, const tp_track_params *p, int n, tp_particle *ptl,
                            const tp_particle *ptl_init, int *nsteps, double *espect,
                            double *espect_escape, double *dcoeff, int *nptl_remain,
                            int nsteps_output, const int *ntraj_accum, tp_particle *ptl_time,
                            int *points)
{
    if (!c || !p) return TP_E_ARG;
    int rc = traj_arm(c, n, nsteps_output, ntraj_accum, ptl_time);
    if (rc) return rc;
    rc = tp_gpu_track_fixed(c, p, n, ptl, ptl_init, nsteps, espect, espect_escape, dcoeff,
                            nptl_remain);
    if (rc != TP_OK && rc != TP_E_NONFINITE) {
        c->traj_nout = 0;
        return rc;
    }
    const int rc2 = traj_collect(c, n, ntraj_accum, ptl_time, points);
    return rc2 ? rc2 : rc;
}

int tp_gpu_track_adaptive_diag_traj(tp_ctx *c, const tp_bs_params *bs, const tp_track_params *p,
                                    int n, tp_particle *ptl, const tp_particle *ptl_init,
                                    int *nsteps, double *espect, double *espect_escape,
                                    double *dcoeff, int *nptl_remain, int nsteps_output,
                                    const int *ntraj_accum, tp_particle *ptl_time, int *points,
                                    tp_bs_stats *stats)
{
    if (!c || !bs || !p || n < 0 || (n > 0 && (!ptl || !ptl_init))) return TP_E_ARG;
    int rc = traj_arm(c, n, nsteps_output, ntraj_accum, ptl_time);
    if (rc) return rc;
    rc = tp_gpu_load_particles(c, n, ptl, ptl_init);
    if (!rc) rc = tp_gpu_run_adaptive_diag(c, bs, p, stats);
    if (!rc) rc = tp_gpu_store_particles(c, ptl, nsteps);
    if (!rc) rc = tp_gpu_read_diagnostics(c, espect, espect_escape, dcoeff, nptl_remain);
    if (rc) {
        c->traj_nout = 0;
        return rc;
    }
    return traj_collect(c, n, ntraj_accum, ptl_time, points);
}

int tp_gpu_track_fixed(tp_ctx *c, const tp_track_params *p, int n, tp_particle *ptl,
                       const tp_particle *ptl_init, int *nsteps, double *espect,
                       double *espect_escape, double *dcoeff, int *nptl_remain)
{
    if (!c || !p || n < 0 || (n > 0 && (!ptl || !ptl_init))) return TP_E_ARG;
    int rc = tp_gpu_load_particles(c, n, ptl, ptl_init);
    if (rc) return rc;
    /* TP_E_NONFINITE (TP_FLAG_CHECK_FINITE) still returns the particles and
     * diagnostics */
    const int run = tp_gpu_run_fixed(c, p);
    if (run != TP_OK && run != TP_E_NONFINITE) return run;
    rc = tp_gpu_store_particles(c, ptl, nsteps);
    if (rc) return rc;
    rc = tp_gpu_read_diagnostics(c, espect, espect_escape, dcoeff, nptl_remain);
    if (rc) return rc;
    return run;
}

} /* extern "C" */
