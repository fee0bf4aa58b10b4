/*
 * tp_fieldline.hip -- Poincare maps of the magnetic field on gfx950.
 *
 * The reference's run_type 2 (src/poincare_map.c:39-258) traces each field
 * line from its start point by classical RK4 on B/|B| (next_location,
 * :268-320), folds every stage point once into the periodic box
 * (field_boundary, :326-352) and records the positions where the line
 * crosses the plane face = start[face] (:166-184).  Lines are independent,
 * so each lane traces one line; the FP64 arithmetic is the reference's, in
 * its order (the library is built with -ffp-contract=off), so crossing
 * points are bitwise those of the CPU oracle.
 *
 * GPU organisation:
 *  - a launch advances every active line by at most `chunk` RK4 steps, so
 *    kernel durations stay bounded however long the lines are; the line
 *    state (position, traced length, next slot) lives in HBM between
 *    launches;
 *  - lines that finish (ncross crossings, or length_max) leave the active
 *    list: each launch appends the lines still running to the next list
 *    (one atomic per wavefront), so later launches run dense wavefronts
 *    instead of waves with idle lanes;
 *  - on the PIC grid the B values of the current cell's eight corners stay
 *    in registers (24 floats) and are reloaded only when a stage point
 *    enters another cell (a line advances step/dx cells per stage); the
 *    cell test compares the six corner indices, and flat node indices are
 *    formed only on a reload.
 *
 * Each line is a serial chain of dependent FP64 operations (four divisions,
 * a square root and the reference's eight float-rounded additions per
 * component per stage), so a wavefront whose lines outlast the others runs
 * at the latency of that chain; lines that run to length_max long after the
 * rest set the run time of a map.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include "tp_ctx.h"

#define PM_CHUNK_DEFAULT 16384
#define PM_CHUNK_MAX (1 << 20)

namespace {

/* B of the eight corners of the current cell.  The reference-gather kinds
 * identify the cell by its six corner indices (key), the textbook kinds by
 * the flat indices of corners 0 and 7 (k0, k7) as the tracker's cell cache
 * does; a cell whose indices were clamped into the array is never reused. */
struct BCache {
    unsigned k0, k7;
    int key[6];
    bool valid;
    float v[24];
};

/* node = ex ey ez bx | by bz 0 0: bx, by, bz are floats 3..5 */
__device__ __forceinline__ void bc_load(const float4 *g, BCache &bc, const unsigned *c)
{
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float *p = (const float *)(g + 2 * (size_t)c[k]) + 3;
        bc.v[3 * k + 0] = p[0];
        bc.v[3 * k + 1] = p[1];
        bc.v[3 * k + 2] = p[2];
    }
}

__device__ __forceinline__ void bc_fill(const float4 *g, BCache &bc, const unsigned *c,
                                        bool clamped)
{
    if (c[0] != bc.k0 || c[7] != bc.k7) {
        bc_load(g, bc, c);
        bc.k0 = clamped ? TP_GC_INVALID : c[0];
        bc.k7 = c[7];
    }
}

/* The reference's trilinear parameters (corner order and x/z-exchanged
 * weights of src/interpolation.c:141-181, the same values as trilinear_ref)
 * with the cell's B values in bc.  Flat node indices are formed only when
 * the cell changes: in 32 bits when every corner index lies inside the
 * grid, otherwise in 64 bits with the reference's clamp into the array. */
template <bool FAST>
__device__ __forceinline__ void pm_ref_cell(const FieldDev &f, const float4 *g, BCache &bc,
                                            double x, double y, double z, double *w)
{
    int i1[3], i2[3];
    double fx, fy, fz;
    ref_cell_index<FAST>(f, x, y, z, i1, i2, fx, fy, fz);
    const bool same = bc.valid && i1[0] == bc.key[0] && i2[0] == bc.key[1] &&
                      i1[1] == bc.key[2] && i2[1] == bc.key[3] && i1[2] == bc.key[4] &&
                      i2[2] == bc.key[5];
    if (!same) {
        unsigned c[8];
        bool cl = false;
        const bool inside = (unsigned)i1[0] < (unsigned)f.nx && (unsigned)i2[0] < (unsigned)f.nx &&
                            (unsigned)i1[1] < (unsigned)f.ny && (unsigned)i2[1] < (unsigned)f.ny &&
                            (unsigned)i1[2] < (unsigned)f.nz && (unsigned)i2[2] < (unsigned)f.nz;
        if (inside) {
            const unsigned sx = (unsigned)f.nx, sxy = (unsigned)f.nx * (unsigned)f.ny;
            const unsigned X1 = i1[0], X2 = i2[0];
            const unsigned Y1 = sx * i1[1], Y2 = sx * i2[1];
            const unsigned Z1 = sxy * i1[2], Z2 = sxy * i2[2];
            c[0] = X1 + Y1 + Z1; c[1] = X1 + Y1 + Z2; c[2] = X1 + Y2 + Z1; c[3] = X2 + Y1 + Z1;
            c[4] = X2 + Y2 + Z1; c[5] = X2 + Y1 + Z2; c[6] = X1 + Y2 + Z2; c[7] = X2 + Y2 + Z2;
        } else {
            const long total = (long)f.nx * f.ny * f.nz;
            const long sx = f.nx, sxy = (long)f.nx * f.ny;
            long a[8];
            a[0] = i1[0] + sx * i1[1] + sxy * i1[2];
            a[1] = i1[0] + sx * i1[1] + sxy * i2[2];
            a[2] = i1[0] + sx * i2[1] + sxy * i1[2];
            a[3] = i2[0] + sx * i1[1] + sxy * i1[2];
            a[4] = i2[0] + sx * i2[1] + sxy * i1[2];
            a[5] = i2[0] + sx * i1[1] + sxy * i2[2];
            a[6] = i1[0] + sx * i2[1] + sxy * i2[2];
            a[7] = i2[0] + sx * i2[1] + sxy * i2[2];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                cl |= (a[k] < 0) | (a[k] >= total);
                c[k] = (unsigned)(a[k] < 0 ? 0 : (a[k] >= total ? total - 1 : a[k]));
            }
        }
        bc_load(g, bc, c);
        bc.key[0] = i1[0]; bc.key[1] = i2[0]; bc.key[2] = i1[1];
        bc.key[3] = i2[1]; bc.key[4] = i1[2]; bc.key[5] = i2[2];
        bc.valid = !cl;
    }
    const double gx = 1.0 - fx, gy = 1.0 - fy, gz = 1.0 - fz;
    w[0] = gx * gy * gz;
    w[1] = fx * gy * gz;
    w[2] = gx * fy * gz;
    w[3] = gx * gy * fz;
    w[4] = gx * fy * fz;
    w[5] = fx * gy * fz;
    w[6] = fx * fy * gz;
    w[7] = fx * fy * fz;
}

/* B at t = 0, as get_magnetic_field returns it (src/emfields.c:76-96) */
template <int FK>
__device__ __forceinline__ void pm_bfield(const FieldDev &f, BCache &bc, double x, double y,
                                          double z, double &bx, double &by, double &bz)
{
    if (FK == FK_GRID_REF || FK == FK_GRID2_REF) {
        double w[8];
        pm_ref_cell<false>(f, FK == FK_GRID_REF ? f.grid : f.grid_b, bc, x, y, z, w);
        if (FK == FK_GRID_REF) {
            /* get_float_bfield_at_point, one slice (src/magnetic_field.c:
             * 603-608): float accumulators rounded after every corner */
            float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                a0 = (float)((double)a0 + (double)bc.v[3 * k + 0] * w[k]);
                a1 = (float)((double)a1 + (double)bc.v[3 * k + 1] * w[k]);
                a2 = (float)((double)a2 + (double)bc.v[3 * k + 2] * w[k]);
            }
            bx = a0; by = a1; bz = a2;
        } else {
            /* two slices (src/magnetic_field.c:609-621) at t = 0: the time
             * fraction is 0, so the blend is the previous slice's double
             * sum (values equal; a zero's sign may differ), rounded to
             * float; the next slice is not read */
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                a0 = a0 + (double)bc.v[3 * k + 0] * w[k];
                a1 = a1 + (double)bc.v[3 * k + 1] * w[k];
                a2 = a2 + (double)bc.v[3 * k + 2] * w[k];
            }
            bx = (double)(float)a0; by = (double)(float)a1; bz = (double)(float)a2;
        }
    } else if (FK == FK_GRID_TB || FK == FK_GRID2_TB) {
        unsigned c[8];
        double w[8];
        trilinear_tb(f.nx, f.ny, f.nz, f.lo, f.hi, f.d, x, y, z, c, w);
        bc_fill(FK == FK_GRID_TB ? f.grid : f.grid_b, bc, c, false);
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            a0 = a0 + (double)bc.v[3 * k + 0] * w[k];
            a1 = a1 + (double)bc.v[3 * k + 1] * w[k];
            a2 = a2 + (double)bc.v[3 * k + 2] * w[k];
        }
        bx = a0; by = a1; bz = a2;
    } else {
        Emf e;
        get_emf64<FK>(f, x, y, z, 0.0, e);
        bx = e.bx; by = e.by; bz = e.bz;
    }
}

template <int FK>
__device__ __forceinline__ void pm_unit(const FieldDev &f, BCache &bc, double x, double y,
                                        double z, double &ux, double &uy, double &uz)
{
    double bx, by, bz;
    pm_bfield<FK>(f, bc, x, y, z, bx, by, bz);
    const double ib = 1.0 / sqrt(bx * bx + by * by + bz * bz);
    ux = bx * ib; uy = by * ib; uz = bz * ib;
}

/* The FP32 form: B in FP32 (float weights and FMA accumulation on the grid,
 * the tracker's FP32 field evaluators otherwise) and the unit vector by the
 * hardware reciprocal square root. */
template <int FK>
__device__ __forceinline__ void pm_unit32(const FieldDev &f, BCache &bc, double x, double y,
                                          double z, float &ux, float &uy, float &uz)
{
    float bx, by, bz;
    if (FK == FK_GRID_REF || FK == FK_GRID2_REF || FK == FK_GRID_TB || FK == FK_GRID2_TB) {
        double w[8];
        const float4 *g = (FK == FK_GRID_REF || FK == FK_GRID_TB) ? f.grid : f.grid_b;
        if (FK == FK_GRID_REF || FK == FK_GRID2_REF) {
            pm_ref_cell<true>(f, g, bc, x, y, z, w);
        } else {
            unsigned c[8];
            trilinear_tb(f.nx, f.ny, f.nz, f.lo, f.hi, f.d, x, y, z, c, w);
            bc_fill(g, bc, c, false);
        }
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float wk = (float)w[k];
            a0 = fmaf(bc.v[3 * k + 0], wk, a0);
            a1 = fmaf(bc.v[3 * k + 1], wk, a1);
            a2 = fmaf(bc.v[3 * k + 2], wk, a2);
        }
        bx = a0; by = a1; bz = a2;
    } else {
        EmfF e;
        get_emf32<FK>(f, x, y, z, 0.0, e);
        bx = e.bx; by = e.by; bz = e.bz;
    }
    const float ib = rsqrtf(fmaf(bz, bz, fmaf(by, by, bx * bx)));
    ux = bx * ib; uy = by * ib; uz = bz * ib;
}

/* FP32 form on a grid: a stage point is the step's base cell (from the
 * folded FP64 position, once per step) plus an FP32 offset q in cell units.
 * Points past the last node fold by n - 1 cells (the grid is endpoint
 * inclusive over the periodic box, so that is one box length), the cell
 * index is kept inside the grid, and the weights follow the field's gather
 * convention (the reference's exchanges the x and z fractions). */
template <int FK>
__device__ __forceinline__ void pm_grid32(const FieldDev &f, BCache &bc, const int *base,
                                          const float *q, float &bx, float &by, float &bz)
{
    constexpr bool REF = FK == FK_GRID_REF || FK == FK_GRID2_REF;
    const float4 *g = (FK == FK_GRID_REF || FK == FK_GRID_TB) ? f.grid : f.grid_b;
    const int n[3] = {f.nx, f.ny, f.nz};
    int i[3];
    float fr[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float fl = floorf(q[a]);
        int k = base[a] + (int)fl;
        fr[a] = q[a] - fl;
        if (k >= n[a] - 1) k -= n[a] - 1;
        if (k < 0) k += n[a] - 1;
        i[a] = min(max(k, 0), n[a] - 2);
    }
    if (!(bc.valid && i[0] == bc.key[0] && i[1] == bc.key[1] && i[2] == bc.key[2])) {
        const unsigned sx = (unsigned)f.nx, sxy = (unsigned)f.nx * (unsigned)f.ny;
        const unsigned b0 = (unsigned)i[0] + sx * (unsigned)i[1] + sxy * (unsigned)i[2];
        const unsigned c[8] = {b0, b0 + sxy, b0 + sx, b0 + 1u, b0 + 1u + sx, b0 + 1u + sxy,
                               b0 + sx + sxy, b0 + 1u + sx + sxy};
        bc_load(g, bc, c);
        bc.key[0] = i[0]; bc.key[1] = i[1]; bc.key[2] = i[2];
        bc.valid = true;
    }
    /* corners in the order (x1y1z1, x1y1z2, x1y2z1, x2y1z1, x2y2z1, x2y1z2,
     * x1y2z2, x2y2z2) */
    const float fx = REF ? fr[2] : fr[0], fy = fr[1], fz = REF ? fr[0] : fr[2];
    const float gx = 1.f - fx, gy = 1.f - fy, gz = 1.f - fz;
    const float w[8] = {gx * gy * gz, gx * gy * fz, gx * fy * gz, fx * gy * gz,
                        fx * fy * gz, fx * gy * fz, gx * fy * fz, fx * fy * fz};
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        a0 = fmaf(bc.v[3 * k + 0], w[k], a0);
        a1 = fmaf(bc.v[3 * k + 1], w[k], a1);
        a2 = fmaf(bc.v[3 * k + 2], w[k], a2);
    }
    bx = a0; by = a1; bz = a2;
}

/* field_boundary (src/poincare_map.c:326-352): one fold per axis */
__device__ __forceinline__ double pm_fold(double v, double lo, double hi, double len)
{
    if (v > hi) v = v - len;
    if (v < lo) v = v + len;
    return v;
}

struct PmapArgs {
    FieldDev f;
    int face, ncross, mode, chunk, stride;
    double step, lmax;
    double lo[3], hi[3], L[3];
    double inv_d[3];             /* 1 / grid spacing (FP32 form) */
    float hd32[3], sd32[3];      /* step / 2 / dx and step / dx (FP32 form) */
    double *px, *py, *pz, *len;  /* line state */
    const double *c0;            /* plane coordinate of each line */
    int *k;                      /* next slot while running; -(next slot) once done */
    long long *nst;              /* RK4 steps taken */
    double *xpos, *ypos, *zpos;  /* records, stride doubles per line */
};

__device__ __forceinline__ double pm_pick(int face, double x, double y, double z)
{
    return face == 0 ? x : (face == 1 ? y : z);
}

/* distance to the plane reduced into [-L/2, L/2] */
__device__ __forceinline__ double pm_offset(double d, double L)
{
    return d - L * rint(d / L);
}

/* Lane gi traces line act[gi] for up to a.chunk RK4 steps. */
template <int FK, int PREC>
__global__ void __launch_bounds__(256) k_pmap(PmapArgs a, int nact, const int *act,
                                              int *act_next, unsigned *cnt)
{
    const int gi = blockIdx.x * 256 + threadIdx.x;
    const bool live = gi < nact;
    const int i = live ? act[gi] : 0;
    int k = live ? a.k[i] : -1;
    if (k > 0) {
        double x = a.px[i], y = a.py[i], z = a.pz[i], len = a.len[i];
        long long ns = a.nst[i];
        const double c0 = a.c0[i];
        const double h = 0.5 * a.step, s6 = a.step / 6;
        BCache bc;
        bc.k0 = TP_GC_INVALID;
        bc.k7 = 0u;
        bc.valid = false;
        for (int it = 0; it < a.chunk; it++) {
            /* next_location (src/poincare_map.c:268-320) */
            const double pre = pm_pick(a.face, x, y, z);
            x = pm_fold(x, a.lo[0], a.hi[0], a.L[0]);
            y = pm_fold(y, a.lo[1], a.hi[1], a.L[1]);
            z = pm_fold(z, a.lo[2], a.hi[2], a.L[2]);
            const double pre_folded = pm_pick(a.face, x, y, z);
            if (PREC == PREC64) {
                /* the four stages in one loop body (one inlined field
                 * gather): stage s evaluates at x + c_s b_{s-1}, c = (0, h,
                 * h, step), and the weighted sum ((b1 + 2 b2) + 2 b3) + b4
                 * is built in the reference's order */
                double ux, uy, uz, sx, sy, sz;
#pragma unroll 1
                for (int s = 0; s < 4; s++) {
                    double tx = x, ty = y, tz = z;
                    if (s > 0) {
                        const double cs = s == 3 ? a.step : h;
                        tx = pm_fold(x + cs * ux, a.lo[0], a.hi[0], a.L[0]);
                        ty = pm_fold(y + cs * uy, a.lo[1], a.hi[1], a.L[1]);
                        tz = pm_fold(z + cs * uz, a.lo[2], a.hi[2], a.L[2]);
                    }
                    pm_unit<FK>(a.f, bc, tx, ty, tz, ux, uy, uz);
                    if (s == 0) {
                        sx = ux; sy = uy; sz = uz;
                    } else if (s < 3) {
                        sx = sx + 2 * ux; sy = sy + 2 * uy; sz = sz + 2 * uz;
                    } else {
                        sx = sx + ux; sy = sy + uy; sz = sz + uz;
                    }
                }
                x = x + s6 * sx;
                y = y + s6 * sy;
                z = z + s6 * sz;
            } else if (FK == FK_GRID_REF || FK == FK_GRID_TB || FK == FK_GRID2_REF ||
                       FK == FK_GRID2_TB) {
                /* FP32 form on a grid: base cell and remainder from the
                 * folded FP64 position once per step, stage offsets,
                 * field, unit vectors and their weighted sum in FP32 */
                int base[3];
                float r0[3];
                const double p3[3] = {x, y, z};
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const double qd = (p3[d] - a.lo[d]) * a.inv_d[d];
                    const double fl = floor(qd);
                    base[d] = tp_floor_to_int(fl);
                    r0[d] = (float)(qd - fl);
                }
                float ux = 0.f, uy = 0.f, uz = 0.f, sx, sy, sz;
#pragma unroll 1
                for (int s = 0; s < 4; s++) {
                    float q[3] = {r0[0], r0[1], r0[2]};
                    if (s > 0) {
                        const float *cs = s == 3 ? a.sd32 : a.hd32;   /* c_s / dx per axis */
                        q[0] = fmaf(cs[0], ux, r0[0]);
                        q[1] = fmaf(cs[1], uy, r0[1]);
                        q[2] = fmaf(cs[2], uz, r0[2]);
                    }
                    float bx, by, bz;
                    pm_grid32<FK>(a.f, bc, base, q, bx, by, bz);
                    const float ib = rsqrtf(fmaf(bz, bz, fmaf(by, by, bx * bx)));
                    ux = bx * ib; uy = by * ib; uz = bz * ib;
                    if (s == 0) {
                        sx = ux; sy = uy; sz = uz;
                    } else {
                        const float wgt = s < 3 ? 2.f : 1.f;
                        sx = fmaf(wgt, ux, sx); sy = fmaf(wgt, uy, sy); sz = fmaf(wgt, uz, sz);
                    }
                }
                x = x + s6 * (double)sx;
                y = y + s6 * (double)sy;
                z = z + s6 * (double)sz;
            } else {
                /* FP32 form, analytic fields: unit vectors and their
                 * weighted sum in FP32, stage points and positions in FP64 */
                float ux, uy, uz, sx, sy, sz;
#pragma unroll 1
                for (int s = 0; s < 4; s++) {
                    double tx = x, ty = y, tz = z;
                    if (s > 0) {
                        const double cs = s == 3 ? a.step : h;
                        tx = pm_fold(x + cs * (double)ux, a.lo[0], a.hi[0], a.L[0]);
                        ty = pm_fold(y + cs * (double)uy, a.lo[1], a.hi[1], a.L[1]);
                        tz = pm_fold(z + cs * (double)uz, a.lo[2], a.hi[2], a.L[2]);
                    }
                    pm_unit32<FK>(a.f, bc, tx, ty, tz, ux, uy, uz);
                    if (s == 0) {
                        sx = ux; sy = uy; sz = uz;
                    } else {
                        const float wgt = s < 3 ? 2.f : 1.f;
                        sx = fmaf(wgt, ux, sx); sy = fmaf(wgt, uy, sy); sz = fmaf(wgt, uz, sz);
                    }
                }
                x = x + s6 * (double)sx;
                y = y + s6 * (double)sy;
                z = z + s6 * (double)sz;
            }
            len = len + a.step;
            ns++;
            const double post = pm_pick(a.face, x, y, z);
            bool hit;
            if (a.mode == TP_PMAP_REF_COMPAT) {
                /* src/poincare_map.c:168-173: the face coordinate before the
                 * step, read before its fold */
                hit = !((pre - c0) * (post - c0) > 0);
            } else {
                const double L = a.L[a.face];
                const double d0 = pm_offset(pre_folded - c0, L), d1 = pm_offset(post - c0, L);
                hit = (d0 * d1 < 0 && fabs(d0 - d1) < 0.5 * L) || (d1 == 0 && d0 != 0);
            }
            if (!(len < a.lmax)) {
                k = -k;
                break;
            }
            if (hit) {
                const size_t o = (size_t)i * a.stride + k;
                a.xpos[o] = x; a.ypos[o] = y; a.zpos[o] = z;
                if (++k > a.ncross) {
                    k = -k;
                    break;
                }
            }
        }
        a.px[i] = x; a.py[i] = y; a.pz[i] = z; a.len[i] = len;
        a.nst[i] = ns;
        a.k[i] = k;
    }
    /* lines still running go to the next active list (one atomic per
     * wavefront) */
    const bool alive = live && k > 0;
    const int lane = threadIdx.x & 63;
    const unsigned long long m = __ballot(alive);
    const unsigned pre = (unsigned)__popcll(m & ((1ull << lane) - 1ull));
    unsigned base = 0u;
    if (lane == 0 && m != 0ull) base = atomicAdd(cnt, (unsigned)__popcll(m));
    base = __shfl(base, 0);
    if (alive) act_next[base + pre] = i;
}

/* ------------------------------------------------------------------ */
/* the tail: four lanes per line                                        */
/* ------------------------------------------------------------------ */
/* Once most lines have finished, the few still running leave a handful of
 * wavefronts, each of which advances at the latency of its own instruction
 * stream.  The tail kernel gives every line a quad of lanes: lane c < 3 of
 * the quad carries component c (position, stage point, cell index along
 * axis c, the gather of B_c over the eight corners, the RK4 sum), lane 3
 * repeats component 2 and writes nothing.  What needs the other axes (the
 * corner weights, the cell test and node indices, |B|) reads them from the
 * quad's lanes through DPP quad permutations and computes it in every lane
 * in the same order as the one-lane kernel, so each component goes through
 * the same IEEE operations and the crossing records are the same bits.  A
 * line's lanes take every branch together (their inputs are equal), so a
 * quad is always whole when it exchanges values. */

/* value of lane K of the quad (quad_perm [K, K, K, K]) */
template <int K>
__device__ __forceinline__ int quad_bcast(int v)
{
    return __builtin_amdgcn_mov_dpp(v, K | (K << 2) | (K << 4) | (K << 6), 0xF, 0xF, false);
}

template <int K>
__device__ __forceinline__ double quad_bcast(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = quad_bcast<K>((int)(unsigned)(b & 0xFFFFFFFFll));
    const int hi = quad_bcast<K>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int K>
__device__ __forceinline__ float quad_bcast(float v)
{
    return __int_as_float(quad_bcast<K>(__float_as_int(v)));
}

/* value of lane k (0..2, the same in every lane) of the quad */
__device__ __forceinline__ double quad_pick(double v, int k)
{
    return k == 0 ? quad_bcast<0>(v) : (k == 1 ? quad_bcast<1>(v) : quad_bcast<2>(v));
}

/* this lane's component of the cell cache: B_c at the eight corners, as
 * doubles, and the six corner indices of the cell */
struct BCacheQ {
    int key[6];
    bool valid;
    double v[8];
};

/* the grid along this lane's axis, in registers */
struct QAxis {
    double lo, hi, d, inv_d;
    int n;
};

__device__ __forceinline__ QAxis pmq_axis(const FieldDev &f, int c)
{
    QAxis g;
    g.lo = c == 0 ? f.lo[0] : (c == 1 ? f.lo[1] : f.lo[2]);
    g.hi = c == 0 ? f.hi[0] : (c == 1 ? f.hi[1] : f.hi[2]);
    g.d = c == 0 ? f.d[0] : (c == 1 ? f.d[1] : f.d[2]);
    g.inv_d = c == 0 ? f.inv_d[0] : (c == 1 ? f.inv_d[1] : f.inv_d[2]);
    g.n = c == 0 ? f.nx : (c == 1 ? f.ny : f.nz);
    return g;
}

/* B_c at a stage point of the reference gather (pm_ref_cell + pm_bfield
 * for FK_GRID_REF, one slice), with t the point's coordinate along this
 * lane's axis c */
__device__ __forceinline__ float pmq_bcomp(const FieldDev &f, const QAxis &g, BCacheQ &bc, int c,
                                           double t)
{
    /* ref_cell_index along axis c */
    const double q = f.exact_inv ? (t - g.lo) * g.inv_d : (t - g.lo) / g.d;
    int ia, ib;
    double fr;
    ref_axis_in(q, ia, ib, fr);
    if (t > g.hi || t < g.lo) {
        /* ref_axis_out: the wrap cell n-1 -> 0 */
        const double v = f.exact_inv ? t * g.inv_d : t / g.d;
        if (t > g.hi) { ia = g.n - 1; ib = 0; fr = v - (g.n - 1); }
        if (t < g.lo) { ia = g.n - 1; ib = 0; fr = v + 1.0; }
    }
    const double fx = quad_bcast<0>(fr), fy = quad_bcast<1>(fr), fz = quad_bcast<2>(fr);
    int i1[3], i2[3];
    i1[0] = quad_bcast<0>(ia); i2[0] = quad_bcast<0>(ib);
    i1[1] = quad_bcast<1>(ia); i2[1] = quad_bcast<1>(ib);
    i1[2] = quad_bcast<2>(ia); i2[2] = quad_bcast<2>(ib);
    const bool same = bc.valid && i1[0] == bc.key[0] && i2[0] == bc.key[1] &&
                      i1[1] == bc.key[2] && i2[1] == bc.key[3] && i1[2] == bc.key[4] &&
                      i2[2] == bc.key[5];
    if (!same) {
        unsigned cn[8];
        bool cl = false;
        const bool inside = (unsigned)i1[0] < (unsigned)f.nx && (unsigned)i2[0] < (unsigned)f.nx &&
                            (unsigned)i1[1] < (unsigned)f.ny && (unsigned)i2[1] < (unsigned)f.ny &&
                            (unsigned)i1[2] < (unsigned)f.nz && (unsigned)i2[2] < (unsigned)f.nz;
        if (inside) {
            const unsigned sx = (unsigned)f.nx, sxy = (unsigned)f.nx * (unsigned)f.ny;
            const unsigned X1 = i1[0], X2 = i2[0];
            const unsigned Y1 = sx * i1[1], Y2 = sx * i2[1];
            const unsigned Z1 = sxy * i1[2], Z2 = sxy * i2[2];
            cn[0] = X1 + Y1 + Z1; cn[1] = X1 + Y1 + Z2; cn[2] = X1 + Y2 + Z1;
            cn[3] = X2 + Y1 + Z1; cn[4] = X2 + Y2 + Z1; cn[5] = X2 + Y1 + Z2;
            cn[6] = X1 + Y2 + Z2; cn[7] = X2 + Y2 + Z2;
        } else {
            const long total = (long)f.nx * f.ny * f.nz;
            const long sx = f.nx, sxy = (long)f.nx * f.ny;
            long a[8];
            a[0] = i1[0] + sx * i1[1] + sxy * i1[2];
            a[1] = i1[0] + sx * i1[1] + sxy * i2[2];
            a[2] = i1[0] + sx * i2[1] + sxy * i1[2];
            a[3] = i2[0] + sx * i1[1] + sxy * i1[2];
            a[4] = i2[0] + sx * i2[1] + sxy * i1[2];
            a[5] = i2[0] + sx * i1[1] + sxy * i2[2];
            a[6] = i1[0] + sx * i2[1] + sxy * i2[2];
            a[7] = i2[0] + sx * i2[1] + sxy * i2[2];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                cl |= (a[k] < 0) | (a[k] >= total);
                cn[k] = (unsigned)(a[k] < 0 ? 0 : (a[k] >= total ? total - 1 : a[k]));
            }
        }
        /* node = ex ey ez bx | by bz 0 0: B_c is float 3 + c */
#pragma unroll
        for (int k = 0; k < 8; k++)
            bc.v[k] = (double)((const float *)(f.grid + 2 * (size_t)cn[k]))[3 + c];
        bc.key[0] = i1[0]; bc.key[1] = i2[0]; bc.key[2] = i1[1];
        bc.key[3] = i2[1]; bc.key[4] = i1[2]; bc.key[5] = i2[2];
        bc.valid = !cl;
    }
    const double gx = 1.0 - fx, gy = 1.0 - fy, gz = 1.0 - fz;
    double w[8];
    w[0] = gx * gy * gz;
    w[1] = fx * gy * gz;
    w[2] = gx * fy * gz;
    w[3] = gx * gy * fz;
    w[4] = gx * fy * fz;
    w[5] = fx * gy * fz;
    w[6] = fx * fy * gz;
    w[7] = fx * fy * fz;
    /* get_float_bfield_at_point (src/magnetic_field.c:603-608): a float
     * accumulator rounded after every corner */
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) acc = (float)((double)acc + bc.v[k] * w[k]);
    return acc;
}

/* the unit vector's component c at a stage point */
__device__ __forceinline__ double pmq_unit(const FieldDev &f, const QAxis &g, BCacheQ &bc, int c,
                                           double t)
{
    const float bcomp = pmq_bcomp(f, g, bc, c, t);
    const double bx = quad_bcast<0>(bcomp), by = quad_bcast<1>(bcomp), bz = quad_bcast<2>(bcomp);
    const double ib = 1.0 / sqrt(bx * bx + by * by + bz * bz);
    return (double)bcomp * ib;
}

/* Quad gi / 4 traces line act[gi / 4] for up to a.chunk RK4 steps (the
 * reference gather of one PIC slice, FP64); the same line state, records
 * and active-list update as k_pmap. */
__global__ void __launch_bounds__(256) k_pmap_quad(PmapArgs a, int nact, const int *act,
                                                   int *act_next, unsigned *cnt)
{
    const int gq = (int)(((unsigned)blockIdx.x * 256u + threadIdx.x) >> 2);
    const int sub = threadIdx.x & 3;
    const int c = sub < 3 ? sub : 2;
    const bool live = gq < nact;
    const int i = live ? act[gq] : 0;
    int k = live ? a.k[i] : -1;
    if (k > 0) {
        const double lo = c == 0 ? a.lo[0] : (c == 1 ? a.lo[1] : a.lo[2]);
        const double hi = c == 0 ? a.hi[0] : (c == 1 ? a.hi[1] : a.hi[2]);
        const double L = c == 0 ? a.L[0] : (c == 1 ? a.L[1] : a.L[2]);
        const QAxis g = pmq_axis(a.f, c);
        double p = c == 0 ? a.px[i] : (c == 1 ? a.py[i] : a.pz[i]);
        double len = a.len[i];
        long long ns = a.nst[i];
        const double c0 = a.c0[i];
        const double h = 0.5 * a.step, s6 = a.step / 6;
        BCacheQ bc;
        bc.valid = false;
#pragma unroll
        for (int j = 0; j < 6; j++) bc.key[j] = 0;
        for (int it = 0; it < a.chunk; it++) {
            /* next_location (src/poincare_map.c:268-320), component c */
            const double pre = quad_pick(p, a.face);
            p = pm_fold(p, lo, hi, L);
            const double pre_folded = quad_pick(p, a.face);
            double u = 0.0, sum = 0.0;
#pragma unroll 1
            for (int st = 0; st < 4; st++) {
                double t = p;
                if (st > 0) t = pm_fold(p + (st == 3 ? a.step : h) * u, lo, hi, L);
                u = pmq_unit(a.f, g, bc, c, t);
                if (st == 0) sum = u;
                else if (st < 3) sum = sum + 2 * u;
                else sum = sum + u;
            }
            p = p + s6 * sum;
            len = len + a.step;
            ns++;
            const double post = quad_pick(p, a.face);
            bool hit;
            if (a.mode == TP_PMAP_REF_COMPAT) {
                hit = !((pre - c0) * (post - c0) > 0);
            } else {
                const double Lf = a.L[a.face];
                const double d0 = pm_offset(pre_folded - c0, Lf), d1 = pm_offset(post - c0, Lf);
                hit = (d0 * d1 < 0 && fabs(d0 - d1) < 0.5 * Lf) || (d1 == 0 && d0 != 0);
            }
            if (!(len < a.lmax)) {
                k = -k;
                break;
            }
            if (hit) {
                const size_t o = (size_t)i * a.stride + k;
                if (sub < 3) (c == 0 ? a.xpos : (c == 1 ? a.ypos : a.zpos))[o] = p;
                if (++k > a.ncross) {
                    k = -k;
                    break;
                }
            }
        }
        if (sub < 3) (c == 0 ? a.px : (c == 1 ? a.py : a.pz))[i] = p;
        if (sub == 0) {
            a.len[i] = len;
            a.nst[i] = ns;
            a.k[i] = k;
        }
    }
    /* lines still running go to the next active list: lane 0 of each quad
     * (one atomic per wavefront) */
    const bool alive = live && k > 0 && sub == 0;
    const int lane = threadIdx.x & 63;
    const unsigned long long m = __ballot(alive);
    const unsigned pre = (unsigned)__popcll(m & ((1ull << lane) - 1ull));
    unsigned base = 0u;
    if (lane == 0 && m != 0ull) base = atomicAdd(cnt, (unsigned)__popcll(m));
    base = __shfl(base, 0);
    if (alive) act_next[base + pre] = i;
}

/* line state and records at the start: slot 0 the start point, the other
 * slots 2*hi (src/poincare_map.c:55-80) */
__global__ void __launch_bounds__(256) k_pmap_init(PmapArgs a, int n, const double *starts,
                                                   double *c0, int *act, unsigned *cnt)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i == 0) { cnt[0] = 0u; cnt[1] = 0u; }
    if (i >= n) return;
    const double x = starts[3 * i], y = starts[3 * i + 1], z = starts[3 * i + 2];
    a.px[i] = x; a.py[i] = y; a.pz[i] = z;
    a.len[i] = 0.0;
    a.nst[i] = 0;
    c0[i] = pm_pick(a.face, x, y, z);
    a.k[i] = (a.ncross >= 1 && 0.0 < a.lmax) ? 1 : -1;
    act[i] = i;
    const size_t o = (size_t)i * a.stride;
    a.xpos[o] = x; a.ypos[o] = y; a.zpos[o] = z;
    const double fx = a.hi[0] * 2, fy = a.hi[1] * 2, fz = a.hi[2] * 2;
    for (int s = 1; s < a.stride; s++) {
        a.xpos[o + s] = fx; a.ypos[o + s] = fy; a.zpos[o + s] = fz;
    }
}

__global__ void k_pmap_zero(unsigned *cnt) { cnt[threadIdx.x] = 0u; }

typedef void (*pmap_fn)(PmapArgs, int, const int *, int *, unsigned *);

template <int FK>
pmap_fn pick_prec(int prec)
{
    return prec == TP_FP64 ? k_pmap<FK, PREC64> : k_pmap<FK, PREC32>;
}

pmap_fn pick_pmap(int fk, int prec)
{
    switch (fk) {
    case FK_TEST: return pick_prec<FK_TEST>(prec);
    case FK_ABC:
    case FK_ABCV: return pick_prec<FK_ABC>(prec);        /* B does not depend on the velocity grid */
    case FK_GRID_REF: return pick_prec<FK_GRID_REF>(prec);
    case FK_GRID_TB: return pick_prec<FK_GRID_TB>(prec);
    case FK_MODES:
    case FK_MLP: return pick_prec<FK_MODES>(prec);       /* the exact mode sum */
    case FK_WLCS: return pick_prec<FK_WLCS>(prec);
    case FK_GRID2_REF: return pick_prec<FK_GRID2_REF>(prec);
    case FK_GRID2_TB: return pick_prec<FK_GRID2_TB>(prec);
    /* system 3: the nodes' B slots hold B as for PIC grids (get_bfield_mhd_
     * test_particle, src/emfields.c:131-143, gathers only B) */
    case FK_MHD_REF: return pick_prec<FK_GRID_REF>(prec);
    case FK_GRIDD_REF: return pick_prec<FK_GRIDD_REF>(prec);
    default: return nullptr;
    }
}

}  // namespace

extern "C" int tp_gpu_poincare_map(tp_ctx *c, const tp_pmap_params *p, int n,
                                   const double *starts, double *xpos, double *ypos,
                                   double *zpos, int *ncrossed, long long *nsteps)
{
    if (!c || !p || n < 0) return TP_E_ARG;
    if (p->face < 0 || p->face > 2 || p->ncross < 0 || p->chunk < 0 || p->chunk > PM_CHUNK_MAX)
        return TP_E_ARG;
    if (p->mode != TP_PMAP_REF_COMPAT && p->mode != TP_PMAP_PERIODIC) return TP_E_ARG;
    if (p->precision != TP_FP64 && p->precision != TP_FP32_MIXED) return TP_E_ARG;
    if (p->lanes != 0 && p->lanes != 1 && p->lanes != 4) return TP_E_ARG;
    /* the traced length grows by step per RK4 step: a finite length_max
     * bounds the steps of every line, and length_max / step < 2^50 keeps
     * len + step > len */
    if (!(p->step > 0.0) || !isfinite(p->step) || !isfinite(p->length_max) ||
        !(p->length_max / p->step < 0x1p50))
        return TP_E_ARG;
    for (int d = 0; d < 3; d++)
        if (!(p->hi[d] > p->lo[d]) || !isfinite(p->hi[d] - p->lo[d])) return TP_E_ARG;
    if (n > 0 && (!starts || !xpos || !ypos || !zpos)) return TP_E_ARG;
    if (!c->have_field) return TP_E_STATE;
    const pmap_fn fn = pick_pmap(field_kind(c), p->precision);
    if (!fn) return TP_E_STATE;
    /* four lanes per line: the reference gather of one PIC slice in FP64
     * (system 3 gathers B the same way) */
    const bool quad_ok = p->precision == TP_FP64 &&
                         (field_kind(c) == FK_GRID_REF || field_kind(c) == FK_MHD_REF);
    const int quad_from = !quad_ok || p->lanes == 1 ? 0 : (p->lanes == 4 ? (n < (1 << 28) ? n : (1 << 28)) : TP_PMAP_QUAD_LINES);
    c->last_ms = 0.0;
    c->last_launches = 0;
    if (n == 0) return TP_OK;
    HIPCHK(hipSetDevice(c->device));

    const int stride = p->ncross + 1;
    const size_t nrec = (size_t)n * stride;
    /* one allocation: 5 doubles + 3 ints + 1 long long per line, records,
     * counters */
    const size_t bytes = sizeof(double) * (size_t)n * 5 + sizeof(long long) * (size_t)n +
                         sizeof(int) * (size_t)n * 3 + sizeof(double) * nrec * 3 + 256;
    char *buf = nullptr;
    if (hipMalloc(&buf, bytes) != hipSuccess) return TP_E_NOMEM;
    double *dv = (double *)buf;
    PmapArgs a;
    memset(&a, 0, sizeof(a));
    a.f = c->f;
    a.face = p->face;
    a.ncross = p->ncross;
    a.mode = p->mode;
    a.chunk = p->chunk > 0 ? p->chunk : PM_CHUNK_DEFAULT;
    a.stride = stride;
    a.step = p->step;
    a.lmax = p->length_max;
    for (int d = 0; d < 3; d++) {
        a.lo[d] = p->lo[d];
        a.hi[d] = p->hi[d];
        a.L[d] = p->hi[d] - p->lo[d];
        a.inv_d[d] = c->f.d[d] != 0.0 ? 1.0 / c->f.d[d] : 0.0;
        a.hd32[d] = (float)(0.5 * p->step * a.inv_d[d]);
        a.sd32[d] = (float)(p->step * a.inv_d[d]);
    }
    a.px = dv;
    a.py = dv + n;
    a.pz = dv + 2 * (size_t)n;
    a.len = dv + 3 * (size_t)n;
    double *c0 = dv + 4 * (size_t)n;
    a.c0 = c0;
    a.xpos = dv + 5 * (size_t)n;
    a.ypos = a.xpos + nrec;
    a.zpos = a.ypos + nrec;
    a.nst = (long long *)(a.zpos + nrec);
    a.k = (int *)(a.nst + n);
    int *act[2] = {a.k + n, a.k + 2 * (size_t)n};
    unsigned *cnt = (unsigned *)(act[1] + n);
    double *dstart = nullptr;

    int rc = TP_OK;
    if (hipMalloc(&dstart, sizeof(double) * 3 * (size_t)n) != hipSuccess) rc = TP_E_NOMEM;
    if (rc == TP_OK && hipMemcpyAsync(dstart, starts, sizeof(double) * 3 * (size_t)n,
                                      hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = TP_E_HIP;
    if (rc == TP_OK) {
        hipLaunchKernelGGL(k_pmap_init, dim3((n + 255) / 256), dim3(256), 0, c->stream, a, n,
                           (const double *)dstart, c0, act[0], cnt);
        if (hipGetLastError() != hipSuccess) rc = TP_E_HIP;
    }
    /* every launch advances each running line by a.chunk steps or ends it */
    const double max_steps = ceil(p->length_max / p->step) + 1.0;
    const long long max_launches = (long long)(max_steps / a.chunk) + 2;
    int nact = n, cur = 0;
    double ms_total = 0.0;
    int launches = 0;
    while (rc == TP_OK && nact > 0) {
        if (launches >= max_launches) { rc = TP_E_HIP; break; }
        unsigned *cn = cnt + (cur ^ 1);
        hipLaunchKernelGGL(k_pmap_zero, dim3(1), dim3(1), 0, c->stream, cn);
        if (hipEventRecord(c->ev0, c->stream) != hipSuccess) { rc = TP_E_HIP; break; }
        if (nact <= quad_from)
            hipLaunchKernelGGL(k_pmap_quad, dim3((4 * nact + 255) / 256), dim3(256), 0,
                               c->stream, a, nact, (const int *)act[cur], act[cur ^ 1], cn);
        else
            hipLaunchKernelGGL(fn, dim3((nact + 255) / 256), dim3(256), 0, c->stream, a, nact,
                               (const int *)act[cur], act[cur ^ 1], cn);
        if (hipGetLastError() != hipSuccess) { rc = TP_E_HIP; break; }
        unsigned h = 0;
        if (hipEventRecord(c->ev1, c->stream) != hipSuccess ||
            hipMemcpyAsync(&h, cn, sizeof(h), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            rc = TP_E_HIP;
            break;
        }
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) ms_total += ms;
        launches++;
        if (h > (unsigned)nact) { rc = TP_E_HIP; break; }
        nact = (int)h;
        cur ^= 1;
    }
    if (rc == TP_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = TP_E_HIP;
    if (rc == TP_OK) {
        if (hipMemcpy(xpos, a.xpos, sizeof(double) * nrec, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(ypos, a.ypos, sizeof(double) * nrec, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(zpos, a.zpos, sizeof(double) * nrec, hipMemcpyDeviceToHost) != hipSuccess)
            rc = TP_E_HIP;
    }
    if (rc == TP_OK && nsteps &&
        hipMemcpy(nsteps, a.nst, sizeof(long long) * (size_t)n, hipMemcpyDeviceToHost) !=
            hipSuccess)
        rc = TP_E_HIP;
    if (rc == TP_OK && ncrossed) {
        if (hipMemcpy(ncrossed, a.k, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost) !=
            hipSuccess) {
            rc = TP_E_HIP;
        } else {
            /* every line has ended: k = -(next slot) */
            for (int i = 0; i < n; i++) ncrossed[i] = -ncrossed[i] - 1;
        }
    }
    (void)hipFree(dstart);
    (void)hipFree(buf);
    c->last_ms = ms_total;
    c->last_launches = launches;
    return rc;
}
