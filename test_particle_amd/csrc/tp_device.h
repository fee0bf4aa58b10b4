/*
 * tp_device.h -- device-side field sources and particle pushers (gfx950).
 *
 * FP64 paths perform the operations of the reference's routines in the
 * reference's order (the library is compiled with -ffp-contract=off, and
 * sin/cos come from tp_math.h), so that results are bitwise those of the
 * CPU oracle.  FP32 paths (TP_FP32_MIXED) evaluate fields and push in
 * float and are compared against FP64 by tolerance.
 */
#ifndef TP_DEVICE_H
#define TP_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "tp_math.h"

#define TP_PI 3.14159265358979323846

/* field kinds, one kernel instantiation each */
enum {
    FK_TEST = 0,      /* B = z-hat */
    FK_ABC = 1,       /* ABC, E = 0 */
    FK_ABCV = 2,      /* ABC, E = B x v(grid) */
    FK_GRID_REF = 3,  /* PIC grid, reference gather */
    FK_GRID_TB = 4,   /* PIC grid, textbook gather */
    FK_MODES = 5,     /* multi-mode turbulence */
    FK_WLCS = 6,      /* wire-loop current systems */
    FK_MLP = 7,       /* learned surrogate of the modes, exact modes as fallback */
    FK_GRID2_REF = 8, /* PIC grid, two time slices, reference gather */
    FK_GRID2_TB = 9,  /* PIC grid, two time slices, textbook gather */
    FK_MHD_REF = 10,  /* MHD grids (system 3): float v and B, E = B x v */
    FK_GRIDD_REF = 11,/* PIC grid of double E and B, reference weights */
};

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

/* A surrogate network on the device (layout in tp_mlp.hip): bfloat16 weight
 * fragments of layers 1-3 in the order the matrix-core kernels read them,
 * FP32 biases, layer 4 and Fourier frequencies, the residual check and the
 * per-particle fallback flags. */
struct MlpNet {
    const bf16x8 *w1f, *w2f, *w3f, *w2b, *w3b;
    const float *b1, *b2, *b3, *w4, *b4;
    const float *q;              /* 128 x 3 frequencies (cycles per unit length) */
    float scale;                 /* output scale */
    float tol;                   /* residual tolerance |dB| / |B| */
    int check;                   /* residual check every `check` steps (0: start only) */
    int *fb;                     /* per particle: 1 = uses the exact field */
    unsigned *nfb;               /* particles switched to the exact field */
};

/* One WLCS system with the trigonometric values read_wlcs precomputes
 * (src/wlcs.c:166-173) and currents scaled by normI (:209-212). */
struct WlcsSys {
    double cur_wr, x_wr, y_wr, z_wr, ct, st, cp, sp, t0;
    double cur_lp, r_lp, x_lp, y_lp, z_lp, ca, sa, cb, sb, omega;
};

struct FieldDev {
    int kind;
    /* ABC */
    double A, B, C, lam, B0;
    /* velocity grid (E = B x v) */
    int vnx, vny, vnz;
    double vlo[3], vhi[3], vd[3];
    const double *vgrid;            /* AoS double3 */
    /* PIC grid, nodes of 8 floats: ex ey ez bx | by bz 0 0 */
    int nx, ny, nz;
    double lo[3], hi[3], d[3];
    double inv_d[3];                /* 1 / d (the FP32 forms' cell index) */
    /* every d is a power of two (normal, with a normal reciprocal): then
     * v / d and v * inv_d are the same double for every v, and the FP64
     * forms multiply instead of dividing */
    int exact_inv;
    /* set by a fixed-step run, not by the field: periodic particles whose
     * box is the grid's own [lo, hi], so particle_bc keeps every position
     * in the box between steps (the FP32 forms' cell index uses it) */
    int in_box;
    const float4 *grid;
    /* two time slices: grid = next slice, grid_b = previous slice, gdt the
     * time between slices */
    const float4 *grid_b;
    double gdt;
    /* cell-major copies of grid and grid_b (null when not built): cell
     * (i, j, k) = the reference gather's eight corners (i1 = i, i2 = i + 1,
     * or 0 after the last node) in its corner order, six floats each, 12
     * float4 per cell, cells indexed like the nodes.  A cell change reads one
     * contiguous 192-byte record instead of eight nodes. */
    const float4 *cells, *cells_b;
    /* PIC grid of doubles (FK_GRIDD_REF): nodes of 8 doubles,
     * ex ey ez bx by bz 0 0, x fastest */
    const double *gridd;
    /* modes: rows kx ky kz ax ay az phi */
    int nmodes;
    const double *modes;
    double b0z;
    /* the same table for FP32: (kx ky kz phi) / 2 pi, ax ay az 0 per mode */
    const float *modes32;
    MlpNet mlp;
    /* WLCS */
    int nwlcs, ibessel, wlmode;
    const WlcsSys *wl;
};

struct Emf {
    double bx, by, bz, ex, ey, ez;
};

struct EmfF {
    float bx, by, bz, ex, ey, ez;
};

/* ------------------------------------------------------------------ */
/* trilinear parameters                                                */
/* ------------------------------------------------------------------ */

__device__ __forceinline__ int tp_floor_to_int(double v)
{
    /* (int) of an out-of-range value is INT_MIN on the reference's x86-64 */
    return (v > -2147483649.0 && v < 2147483648.0) ? (int)v : (int)0x80000000;
}

/* grid_indices + get_trilinar_parameters (src/interpolation.c:36-181):
 * corner order and x/z-exchanged weights of the reference; flat corner
 * index clamped into the array. */
__device__ __forceinline__ void trilinear_ref(int nx, int ny, int nz,
        const double *lo, const double *hi, const double *dd,
        double x, double y, double z, unsigned *c, double *w, bool *clamped = nullptr)
{
    double qx = (x - lo[0]) / dd[0];
    double qy = (y - lo[1]) / dd[1];
    double qz = (z - lo[2]) / dd[2];
    int ix1 = tp_floor_to_int(floor(qx));
    int iy1 = tp_floor_to_int(floor(qy));
    int iz1 = tp_floor_to_int(floor(qz));
    int ix2 = ix1 + 1, iy2 = iy1 + 1, iz2 = iz1 + 1;
    double fx = qx - ix1, fy = qy - iy1, fz = qz - iz1;
    if (x > hi[0]) { ix1 = nx - 1; ix2 = 0; fx = x / dd[0] - (nx - 1); }
    if (y > hi[1]) { iy1 = ny - 1; iy2 = 0; fy = y / dd[1] - (ny - 1); }
    if (z > hi[2]) { iz1 = nz - 1; iz2 = 0; fz = z / dd[2] - (nz - 1); }
    if (x < lo[0]) { ix1 = nx - 1; ix2 = 0; fx = x / dd[0] + 1.0; }
    if (y < lo[1]) { iy1 = ny - 1; iy2 = 0; fy = y / dd[1] + 1.0; }
    if (z < lo[2]) { iz1 = nz - 1; iz2 = 0; fz = z / dd[2] + 1.0; }
    long total = (long)nx * ny * nz;
    long sx = nx, sxy = (long)nx * ny;
    long a[8];
    a[0] = ix1 + sx * iy1 + sxy * iz1;
    a[1] = ix1 + sx * iy1 + sxy * iz2;
    a[2] = ix1 + sx * iy2 + sxy * iz1;
    a[3] = ix2 + sx * iy1 + sxy * iz1;
    a[4] = ix2 + sx * iy2 + sxy * iz1;
    a[5] = ix2 + sx * iy1 + sxy * iz2;
    a[6] = ix1 + sx * iy2 + sxy * iz2;
    a[7] = ix2 + sx * iy2 + sxy * iz2;
    bool cl = false;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        cl |= (a[k] < 0) | (a[k] >= total);
        c[k] = (unsigned)(a[k] < 0 ? 0 : (a[k] >= total ? total - 1 : a[k]));
    }
    if (clamped) *clamped = cl;
    double gx = 1.0 - fx, gy = 1.0 - fy, gz = 1.0 - fz;
    w[0] = gx * gy * gz;
    w[1] = fx * gy * gz;
    w[2] = gx * fy * gz;
    w[3] = gx * gy * fz;
    w[4] = gx * fy * fz;
    w[5] = fx * gy * fz;
    w[6] = fx * fy * gz;
    w[7] = fx * fy * fz;
}

/* textbook trilinear: fold by the period, corner ix+1 weighted by fx */
__device__ __forceinline__ void trilinear_tb(int nx, int ny, int nz,
        const double *lo, const double *hi, const double *dd,
        double x, double y, double z, unsigned *c, double *w)
{
    double p[3] = {x, y, z};
    int n[3] = {nx, ny, nz};
    int i1[3];
    double f[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double L = hi[a] - lo[a];
        double q = p[a] - lo[a];
        q = q - L * floor(q / L);
        double v = floor(q / dd[a]);
        int i = (v >= 0.0 && v <= n[a] - 2.0) ? (int)v : (v < 0.0 ? 0 : n[a] - 2);
        if (!(v == v)) i = 0;
        i1[a] = i;
        f[a] = q / dd[a] - i;
    }
    unsigned sx = nx, sxy = (unsigned)nx * ny;
    unsigned b = i1[0] + sx * i1[1] + sxy * i1[2];
    c[0] = b;
    c[1] = b + sxy;
    c[2] = b + sx;
    c[3] = b + 1;
    c[4] = b + 1 + sx;
    c[5] = b + 1 + sxy;
    c[6] = b + sx + sxy;
    c[7] = b + 1 + sx + sxy;
    double fx = f[0], fy = f[1], fz = f[2];
    double gx = 1.0 - fx, gy = 1.0 - fy, gz = 1.0 - fz;
    w[0] = gx * gy * gz;
    w[1] = gx * gy * fz;
    w[2] = gx * fy * gz;
    w[3] = fx * gy * gz;
    w[4] = fx * fy * gz;
    w[5] = fx * gy * fz;
    w[6] = gx * fy * fz;
    w[7] = fx * fy * fz;
}

/* ------------------------------------------------------------------ */
/* register-resident cell cache for the grid gather                    */
/* ------------------------------------------------------------------ */
/* A particle moves ~0.06 cells per step, so it gathers from the same eight
 * corners for ~10 steps in a row.  The eight corners' six components (48
 * floats) stay in VGPRs and are reloaded only when the corner set changes.
 * The reference gather identifies the set by its six per-axis corner
 * indices (the flat node indices are formed only on a reload); the textbook
 * gather by the flat indices of corners 0 and 7: c7 - c0 = dx + nx dy +
 * nx ny dz with each d either 1 or 1 - n (a wrap cell), and these eight
 * sums are distinct, so (c0, c7) fixes all eight corners.  Cells whose
 * indices were clamped are never reused.  The gathered values are the same
 * floats either way, so results are bit-identical to the uncached gather. */
#define TP_GC_INVALID 0xFFFFFFFFu

/* NS time slices: v[48 s + 6 k + component] of corner k in slice s (slice 0
 * = grid, slice 1 = grid_b) */
template <int NS>
struct GridCacheN {
    unsigned k0, k7;   /* textbook gather: flat indices of corners 0 and 7 */
    int key[6];        /* reference gather: the six corner indices */
    bool valid;
    float v[48 * NS];
};
typedef GridCacheN<1> GridCache;

/* the cache a field kind's step kernel carries */
template <int FK>
struct CacheFor {
    typedef GridCacheN<(FK == FK_GRID2_REF || FK == FK_GRID2_TB) ? 2 : 1> type;
};

template <int NS>
__device__ __forceinline__ void gc_init(GridCacheN<NS> &gc)
{
    gc.k0 = TP_GC_INVALID;
    gc.k7 = 0u;
    gc.valid = false;
}

template <int NS>
__device__ __forceinline__ void gc_fill(const FieldDev &f, GridCacheN<NS> &gc,
                                        const unsigned *c, bool clamped)
{
    if (c[0] != gc.k0 || c[7] != gc.k7) {
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const float4 *g = s == 0 ? f.grid : f.grid_b;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const float4 p = g[2 * (size_t)c[k]];
                const float2 q = *(const float2 *)(g + 2 * (size_t)c[k] + 1);
                float *v = gc.v + 48 * s + 6 * k;
                v[0] = p.x; v[1] = p.y; v[2] = p.z;
                v[3] = p.w; v[4] = q.x; v[5] = q.y;
            }
        }
        gc.k0 = clamped ? TP_GC_INVALID : c[0];
        gc.k7 = c[7];
    }
}

template <int NS>
__device__ __forceinline__ void gc_load(const FieldDev &f, GridCacheN<NS> &gc, const unsigned *c)
{
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const float4 *g = s == 0 ? f.grid : f.grid_b;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float4 p = g[2 * (size_t)c[k]];
            const float2 q = *(const float2 *)(g + 2 * (size_t)c[k] + 1);
            float *v = gc.v + 48 * s + 6 * k;
            v[0] = p.x; v[1] = p.y; v[2] = p.z;
            v[3] = p.w; v[4] = q.x; v[5] = q.y;
        }
    }
}

/* grid_indices / get_trilinar_parameters (src/interpolation.c:36-105):
 * the point in cell units, (v - lo) / d on each axis.  FAST (the FP32 forms
 * only) multiplies by 1/d; the FP64 forms divide as the reference does,
 * except on a grid whose spacings are powers of two, where the product with
 * the exact reciprocal is the same double.  One uniform branch for the
 * three axes keeps their divisions independent of each other. */
template <bool FAST>
__device__ __forceinline__ void ref_cell_coords(const FieldDev &f, double x, double y, double z,
                                                double *q)
{
    if (FAST || f.exact_inv) {
        q[0] = (x - f.lo[0]) * f.inv_d[0];
        q[1] = (y - f.lo[1]) * f.inv_d[1];
        q[2] = (z - f.lo[2]) * f.inv_d[2];
    } else {
        q[0] = (x - f.lo[0]) / f.d[0];
        q[1] = (y - f.lo[1]) / f.d[1];
        q[2] = (z - f.lo[2]) / f.d[2];
    }
}

/* along one axis: lower and upper corner and the fraction of the point at
 * q cell units */
__device__ __forceinline__ void ref_axis_in(double q, int &i1, int &i2, double &fr)
{
    i1 = tp_floor_to_int(floor(q));
    i2 = i1 + 1;
    fr = q - i1;
}

/* the wrap cell n-1 -> 0, for a point beyond either end of [lo, hi] */
template <bool FAST>
__device__ __forceinline__ void ref_axis_out(const FieldDev &f, int a, double x, int n, int &i1,
                                             int &i2, double &fr)
{
    const bool mul = FAST || f.exact_inv;
    if (x > f.hi[a]) { i1 = n - 1; i2 = 0; fr = (mul ? x * f.inv_d[a] : x / f.d[a]) - (n - 1); }
    if (x < f.lo[a]) { i1 = n - 1; i2 = 0; fr = (mul ? x * f.inv_d[a] : x / f.d[a]) + 1.0; }
}

/* the reference's corner indices and fractions of the point (x, y, z):
 * the branch for points outside the box is skipped by a wave none of whose
 * points is outside (the usual case) */
template <bool FAST>
__device__ __forceinline__ void ref_cell_index(const FieldDev &f, double x, double y, double z,
                                               int *i1, int *i2, double &fx, double &fy,
                                               double &fz)
{
    double q[3];
    ref_cell_coords<FAST>(f, x, y, z, q);
    ref_axis_in(q[0], i1[0], i2[0], fx);
    ref_axis_in(q[1], i1[1], i2[1], fy);
    ref_axis_in(q[2], i1[2], i2[2], fz);
    if (x > f.hi[0] || x < f.lo[0] || y > f.hi[1] || y < f.lo[1] || z > f.hi[2] ||
        z < f.lo[2]) {
        ref_axis_out<FAST>(f, 0, x, f.nx, i1[0], i2[0], fx);
        ref_axis_out<FAST>(f, 1, y, f.ny, i1[1], i2[1], fy);
        ref_axis_out<FAST>(f, 2, z, f.nz, i1[2], i2[2], fz);
    }
}

/* The reference gather's parameters (the values of trilinear_ref: corner
 * order and x/z-exchanged weights of src/interpolation.c:141-181) with the
 * cell's nodes cached in gc.  The cell is identified by its six corner
 * indices, and flat node indices are formed only when it changes: in 32
 * bits when every corner index lies inside the grid, otherwise in 64 bits
 * with the clamp into the array (a clamped cell is never reused). */
/* One axis of a point that particle_bc has folded into [lo, hi]: the lower
 * corner is the integer part of the cell coordinate q = (v - lo) / d and the
 * fraction q - floor(q), which is exact (v_fract); q is clamped into
 * [0, n - 1] first so that the conversion is defined for any v (a NaN
 * gives corner 0 and a NaN fraction).  For 0 <= q <= n - 1 these are the
 * values ref_axis_in gives, and no point in the box needs the wrap cell. */
__device__ __forceinline__ void ref_axis_box(double v, double lo, double inv_d, int n, int &i1,
                                             double &fr)
{
    const double q = (v - lo) * inv_d;
    fr = __builtin_amdgcn_fract(q);
    i1 = (int)fmin(fmax(q, 0.0), (double)(n - 1));
}

template <bool FAST, int NS>
__device__ __forceinline__ void ref_cell_cached(const FieldDev &f, GridCacheN<NS> &gc,
                                                double x, double y, double z, double &fx,
                                                double &fy, double &fz)
{
    int i1[3], i2[3];
    bool same;
    if (FAST && f.in_box) {
        /* the FP32 forms with periodic particles in the grid's box: the
         * corners i2 = i1 + 1 follow from i1, which alone keys the cache
         * (the launch's cache never holds a wrap cell) */
        ref_axis_box(x, f.lo[0], f.inv_d[0], f.nx, i1[0], fx);
        ref_axis_box(y, f.lo[1], f.inv_d[1], f.ny, i1[1], fy);
        ref_axis_box(z, f.lo[2], f.inv_d[2], f.nz, i1[2], fz);
        i2[0] = i1[0] + 1; i2[1] = i1[1] + 1; i2[2] = i1[2] + 1;
        same = gc.valid && i1[0] == gc.key[0] && i1[1] == gc.key[2] && i1[2] == gc.key[4];
    } else {
        ref_cell_index<FAST>(f, x, y, z, i1, i2, fx, fy, fz);
        same = gc.valid && i1[0] == gc.key[0] && i2[0] == gc.key[1] && i1[1] == gc.key[2] &&
               i2[1] == gc.key[3] && i1[2] == gc.key[4] && i2[2] == gc.key[5];
    }
    if (!same) {
        unsigned c[8];
        bool cl = false;
        const bool inside = (unsigned)i1[0] < (unsigned)f.nx && (unsigned)i2[0] < (unsigned)f.nx &&
                            (unsigned)i1[1] < (unsigned)f.ny && (unsigned)i2[1] < (unsigned)f.ny &&
                            (unsigned)i1[2] < (unsigned)f.nz && (unsigned)i2[2] < (unsigned)f.nz;
        if (inside && f.cells) {
            /* inside the grid the lower corner fixes the cell (i2 = i1 + 1,
             * or 0 for the wrap cell): one record */
            const unsigned cid = (unsigned)i1[0] +
                                 (unsigned)f.nx * ((unsigned)i1[1] + (unsigned)f.ny * (unsigned)i1[2]);
#pragma unroll
            for (int sl = 0; sl < NS; sl++) {
                const float4 *r = (sl == 0 ? f.cells : f.cells_b) + 12 * (size_t)cid;
#pragma unroll
                for (int q = 0; q < 12; q++) {
                    const float4 v = r[q];
                    float *o = gc.v + 48 * sl + 4 * q;
                    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
                }
            }
            gc.key[0] = i1[0]; gc.key[1] = i2[0]; gc.key[2] = i1[1];
            gc.key[3] = i2[1]; gc.key[4] = i1[2]; gc.key[5] = i2[2];
            gc.valid = true;
            return;
        }
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
        gc_load(f, gc, c);
        gc.key[0] = i1[0]; gc.key[1] = i2[0]; gc.key[2] = i1[1];
        gc.key[3] = i2[1]; gc.key[4] = i1[2]; gc.key[5] = i2[2];
        gc.valid = !cl;
    }
}

/* the reference's corner weights (x and z fractions exchanged,
 * src/interpolation.c:173-180), in T */
template <typename T>
__device__ __forceinline__ void ref_weights(T fx, T fy, T fz, T *w)
{
    const T gx = T(1) - fx, gy = T(1) - fy, gz = T(1) - fz;
    w[0] = gx * gy * gz;
    w[1] = fx * gy * gz;
    w[2] = gx * fy * gz;
    w[3] = gx * gy * fz;
    w[4] = gx * fy * fz;
    w[5] = fx * gy * fz;
    w[6] = fx * fy * gz;
    w[7] = fx * fy * fz;
}

/* the reference gather's sums over the cached corners with FP64 weights w:
 * one slice with float accumulators rounded after every corner
 * (src/magnetic_field.c:603-608) */
__device__ __forceinline__ void ref_sum64(const GridCacheN<1> &gc, const double *w, Emf &o)
{
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f, a4 = 0.f, a5 = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        a0 = (float)((double)a0 + (double)gc.v[6 * k + 0] * w[k]);
        a1 = (float)((double)a1 + (double)gc.v[6 * k + 1] * w[k]);
        a2 = (float)((double)a2 + (double)gc.v[6 * k + 2] * w[k]);
        a3 = (float)((double)a3 + (double)gc.v[6 * k + 3] * w[k]);
        a4 = (float)((double)a4 + (double)gc.v[6 * k + 4] * w[k]);
        a5 = (float)((double)a5 + (double)gc.v[6 * k + 5] * w[k]);
    }
    o.ex = a0; o.ey = a1; o.ez = a2; o.bx = a3; o.by = a4; o.bz = a5;
}

/* two slices: sums in double, blended dt * next + (1 - dt) * previous,
 * rounded to float for the reference gather (src/magnetic_field.c:609-621) */
__device__ __forceinline__ void ref_sum64(const GridCacheN<2> &gc, const double *w, double gdt,
                                          double t, bool round_float, Emf &o)
{
    double a[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, b[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 8; k++)
#pragma unroll
        for (int m = 0; m < 6; m++) {
            a[m] = a[m] + (double)gc.v[6 * k + m] * w[k];
            b[m] = b[m] + (double)gc.v[48 + 6 * k + m] * w[k];
        }
    const double r = t / gdt;
    const double dtf = r - (double)tp_floor_to_int(floor(r));
    double v[6];
#pragma unroll
    for (int m = 0; m < 6; m++) {
        v[m] = dtf * a[m] + (1.0 - dtf) * b[m];
        if (round_float) v[m] = (double)(float)v[m];
    }
    o.ex = v[0]; o.ey = v[1]; o.ez = v[2];
    o.bx = v[3]; o.by = v[4]; o.bz = v[5];
}

/* two time slices through the cache: the same sums, blend and rounding as
 * get_emf64<FK_GRID2_*> */
template <int FK>
__device__ __forceinline__ void grid2_emf64_cached(const FieldDev &f, GridCacheN<2> &gc,
                                                   double x, double y, double z, double t,
                                                   Emf &o)
{
    double w[8];
    if (FK == FK_GRID2_REF) {
        double fx, fy, fz;
        ref_cell_cached<false, 2>(f, gc, x, y, z, fx, fy, fz);
        ref_weights<double>(fx, fy, fz, w);
    } else {
        unsigned c[8];
        trilinear_tb(f.nx, f.ny, f.nz, f.lo, f.hi, f.d, x, y, z, c, w);
        gc_fill(f, gc, c, false);
    }
    ref_sum64(gc, w, f.gdt, t, FK == FK_GRID2_REF, o);
}

template <int FK>
__device__ __forceinline__ void grid_emf64_cached(const FieldDev &f, GridCacheN<1> &gc, double x,
                                                  double y, double z, Emf &o)
{
    double w[8];
    if (FK == FK_GRID_REF) {
        double fx, fy, fz;
        ref_cell_cached<false, 1>(f, gc, x, y, z, fx, fy, fz);
        ref_weights<double>(fx, fy, fz, w);
        ref_sum64(gc, w, o);
    } else {
        unsigned c[8];
        trilinear_tb(f.nx, f.ny, f.nz, f.lo, f.hi, f.d, x, y, z, c, w);
        gc_fill(f, gc, c, false);
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0, a5 = 0.0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            a0 = a0 + (double)gc.v[6 * k + 0] * w[k];
            a1 = a1 + (double)gc.v[6 * k + 1] * w[k];
            a2 = a2 + (double)gc.v[6 * k + 2] * w[k];
            a3 = a3 + (double)gc.v[6 * k + 3] * w[k];
            a4 = a4 + (double)gc.v[6 * k + 4] * w[k];
            a5 = a5 + (double)gc.v[6 * k + 5] * w[k];
        }
        o.ex = a0; o.ey = a1; o.ez = a2; o.bx = a3; o.by = a4; o.bz = a5;
    }
}


template <int FK>
__device__ __forceinline__ void grid_emf32_cached(const FieldDev &f, GridCacheN<1> &gc, double x,
                                                  double y, double z, EmfF &o)
{
    float w[8];
    if (FK == FK_GRID_REF) {
        /* FP32 form: cell index by 1/dx, weights in float */
        double fx, fy, fz;
        ref_cell_cached<true, 1>(f, gc, x, y, z, fx, fy, fz);
        ref_weights<float>((float)fx, (float)fy, (float)fz, w);
    } else {
        unsigned c[8];
        double wd[8];
        trilinear_tb(f.nx, f.ny, f.nz, f.lo, f.hi, f.d, x, y, z, c, wd);
        gc_fill(f, gc, c, false);
#pragma unroll
        for (int k = 0; k < 8; k++) w[k] = (float)wd[k];
    }
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f, a4 = 0.f, a5 = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float wk = w[k];
        a0 = fmaf(gc.v[6 * k + 0], wk, a0); a1 = fmaf(gc.v[6 * k + 1], wk, a1);
        a2 = fmaf(gc.v[6 * k + 2], wk, a2); a3 = fmaf(gc.v[6 * k + 3], wk, a3);
        a4 = fmaf(gc.v[6 * k + 4], wk, a4); a5 = fmaf(gc.v[6 * k + 5], wk, a5);
    }
    o.ex = a0; o.ey = a1; o.ez = a2; o.bx = a3; o.by = a4; o.bz = a5;
}

/* two time slices in the FP32 form: each slice summed over the corners in
 * float, blended dt * next + (1 - dt) * previous in float (the FP64 form
 * sums and blends in double, as the reference does) */
template <int FK>
__device__ __forceinline__ void grid2_emf32_cached(const FieldDev &f, GridCacheN<2> &gc, double x,
                                                   double y, double z, double t, EmfF &o)
{
    float w[8];
    if (FK == FK_GRID2_REF) {
        double fx, fy, fz;
        ref_cell_cached<true>(f, gc, x, y, z, fx, fy, fz);
        ref_weights<float>((float)fx, (float)fy, (float)fz, w);
    } else {
        unsigned c[8];
        double wd[8];
        trilinear_tb(f.nx, f.ny, f.nz, f.lo, f.hi, f.d, x, y, z, c, wd);
        gc_fill(f, gc, c, false);
#pragma unroll
        for (int k = 0; k < 8; k++) w[k] = (float)wd[k];
    }
    float a[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, b[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 8; k++)
#pragma unroll
        for (int m = 0; m < 6; m++) {
            a[m] = fmaf(gc.v[6 * k + m], w[k], a[m]);
            b[m] = fmaf(gc.v[48 + 6 * k + m], w[k], b[m]);
        }
    const double r = t / f.gdt;
    const float dtf = (float)(r - floor(r));
    const float dtg = 1.0f - dtf;
    o.ex = dtf * a[0] + dtg * b[0]; o.ey = dtf * a[1] + dtg * b[1]; o.ez = dtf * a[2] + dtg * b[2];
    o.bx = dtf * a[3] + dtg * b[3]; o.by = dtf * a[4] + dtg * b[4]; o.bz = dtf * a[5] + dtg * b[5];
}

/* ------------------------------------------------------------------ */
/* WLCS (src/wlcs.c:281-449, 814-903)                                  */
/* ------------------------------------------------------------------ */
#define TP_WL_L0 6.96E8
#define TP_WL_C0 3.0E10
#define TP_WL_I0 4.0E8

/* complete elliptic integrals by the arithmetic-geometric mean */
__device__ __forceinline__ void tp_ellint_agm(double k, double *E, double *K)
{
    double a = 1.0, b = sqrt(1.0 - k * k), c = k;
    double sum = 0.5 * c * c, pw = 0.5;
    for (int i = 0; i < 40; i++) {
        double an = 0.5 * (a + b);
        double bn = sqrt(a * b);
        c = 0.5 * (a - b);
        pw *= 2.0;
        sum += pw * c * c;
        a = an;
        b = bn;
        if (fabs(c) < 1e-17 * a) break;
    }
    double KK = TP_PI / (2.0 * a);
    *K = KK;
    *E = KK * (1.0 - sum);
}

/* the reference's ellint, whose m >= 0.9 branch takes K(m-0.9) for
 * K(1-m) (src/wlcs.c:741-806) */
__device__ __forceinline__ void tp_ellint_ref(double k, double *E, double *K)
{
    double m = k * k;
    if (!(m >= 0.9)) {
        tp_ellint_agm(k, E, K);
        return;
    }
    double Ec, Kc, Em, Km, Ew, Kw;
    tp_ellint_agm(sqrt(1.0 - m), &Ec, &Kc);
    tp_ellint_agm(k, &Em, &Km);
    tp_ellint_agm(sqrt(m - 0.9), &Ew, &Kw);
    double Kref = (TP_PI * Km / Kc) * (Kw / TP_PI);
    *K = Kref;
    *E = 1.0 / Kw * (TP_PI / 2.0 + Kref * (Kc - Ec));
}

__device__ __forceinline__ void wl_transform(double *ax, double *ay, double *az,
        double c1, double s1, double c2, double s2)
{
    double ax1 = c2 * (*ax) + s2 * (*ay);
    double ay1 = -s2 * (*ax) + c2 * (*ay);
    double az1 = *az;
    *ax = c1 * ax1 - s1 * az1;
    *ay = ay1;
    *az = s1 * ax1 + c1 * az1;
}

__device__ __forceinline__ void wl_reverse(double *ax, double *ay, double *az,
        double c1, double s1, double c2, double s2)
{
    double ax1 = c1 * (*ax) + s1 * (*az);
    double ay1 = *ay;
    double az1 = -s1 * (*ax) + c1 * (*az);
    *ax = c2 * ax1 - s2 * ay1;
    *ay = s2 * ax1 + c2 * ay1;
    *az = az1;
}

__device__ __noinline__ void emf_wlcs(const FieldDev &f, double x, double y,
                                      double z, double t, Emf &o)
{
    double Bx = 0.0, By = 0.0, Bz = 0.0, Ex = 0.0, Ey = 0.0, Ez = 0.0;
    for (int i = 0; i < f.nwlcs; i++) {
        const WlcsSys s = f.wl[i];
        double xp = x - s.x_wr, yp = y - s.y_wr, zp = z - s.z_wr;
        wl_transform(&xp, &yp, &zp, s.ct, s.st, s.cp, s.sp);
        if (f.ibessel == 0) {
            /* getWireB, src/wlcs.c:361-401 */
            double rho = sqrt(xp * xp + yp * yp);
            if (rho < 0.1) rho = 0.1;
            double cphi = xp / rho;
            double phi0 = (yp >= 0.0) ? acos(cphi) : 2.0 * TP_PI - acos(cphi);
            double bphi = 1.0 / 5.0 * (s.cur_wr * TP_WL_I0) / (rho * TP_WL_L0);
            double sn, cs;
            tp_sincos(phi0 + TP_PI / 2.0, &sn, &cs);
            double bx = bphi * cs, by = bphi * sn, bz = 0.0;
            wl_reverse(&bx, &by, &bz, s.ct, s.st, s.cp, s.sp);
            Bx += bx; By += by; Bz += bz;
        } else {
            /* getWire_Bessel, src/wlcs.c:281-355 (float Bessel functions) */
            double t1 = s.t0 + t, om = s.omega;
            double rho = sqrt(xp * xp + yp * yp);
            double x0 = om * rho * TP_WL_L0 / TP_WL_C0;
            float x0f = (float)x0;
            double bj0 = j0f(x0f), bj1 = j1f(x0f), by0 = y0f(x0f), by1 = y1f(x0f);
            double const0 = (s.cur_wr * TP_WL_I0) / (5.0 * rho * TP_WL_L0) * TP_PI / 2.0;
            double sw, cw;
            tp_sincos(om * t1, &sw, &cw);
            double wb = const0 * x0 * (-by1 * cw + bj1 * sw);
            double phi1;
            if (rho == 0.0) phi1 = TP_PI / 2.0;
            else if (yp >= 0.0) phi1 = acos(xp / rho) + TP_PI / 2.0;
            else phi1 = 2.0 * TP_PI - acos(xp / rho) + TP_PI / 2.0;
            double sp1, cp1;
            tp_sincos(phi1, &sp1, &cp1);
            double bx = wb * cp1, by = wb * sp1, bz = 0.0;
            wl_reverse(&bx, &by, &bz, s.ct, s.st, s.cp, s.sp);
            double we = -const0 * x0 * (by0 * sw + bj0 * cw);
            double ex = 0.0, ey = 0.0, ez = we;
            wl_reverse(&ex, &ey, &ez, s.ct, s.st, s.cp, s.sp);
            Bx += bx; Ex += ex;
            By += by; Ey += ey;
            Bz += bz; Ez += ez;
        }
        /* getLoopB, src/wlcs.c:406-449 */
        double lx = (x - s.x_lp) * TP_WL_L0;
        double ly = (y - s.y_lp) * TP_WL_L0;
        double lz = (z - s.z_lp) * TP_WL_L0;
        double al = s.r_lp * TP_WL_L0;
        double curI = s.cur_lp * TP_WL_I0;
        wl_transform(&lx, &ly, &lz, s.cb, s.sb, s.ca, s.sa);
        double Irho = 0.2 * curI;
        double rxy = sqrt(lx * lx + ly * ly);
        double al2 = al * al;
        double r2 = rxy * rxy + lz * lz;
        double irho1 = 1 / sqrt((al + rxy) * (al + rxy) + lz * lz);
        double irho2 = 1 / ((al - rxy) * (al - rxy) + lz * lz);
        double k = sqrt(4.0 * al * rxy) * irho1;
        double elle, ellk;
        if (f.wlmode == 0) tp_ellint_ref(k, &elle, &ellk);
        else tp_ellint_agm(k, &elle, &ellk);
        double Brho = Irho * lz * irho1 * (-ellk + (al2 + r2) * elle * irho2) / rxy;
        if (rxy == 0.0) Brho = 0.0;
        double Bzl = Irho * irho1 * (ellk - (r2 - al2) * elle * irho2);
        Brho /= rxy;
        double bx = Brho * lx, by = Brho * ly, bz = Bzl;
        wl_reverse(&bx, &by, &bz, s.cb, s.sb, s.ca, s.sa);
        Bx += bx; By += by; Bz += bz;
    }
    o.bx = Bx; o.by = By; o.bz = Bz + 0.0;
    o.ex = Ex; o.ey = Ey; o.ez = Ez;
}

/* getemf_mhd_test_particle, src/emfields.c:107-120: the velocity is
 * gathered like E (float accumulators, src/velocity_field.c:593-622) and
 * held in the E slots; E = B x v is formed in double from the float B and v */
__device__ __forceinline__ void mhd_e64(Emf &o)
{
    const double vx = o.ex, vy = o.ey, vz = o.ez;
    o.ex = o.by * vz - o.bz * vy;
    o.ey = o.bz * vx - o.bx * vz;
    o.ez = o.bx * vy - o.by * vx;
}

/* the same in the FP32 form */
__device__ __forceinline__ void mhd_e32(EmfF &o)
{
    const float vx = o.ex, vy = o.ey, vz = o.ez;
    o.ex = o.by * vz - o.bz * vy;
    o.ey = o.bz * vx - o.bx * vz;
    o.ez = o.bx * vy - o.by * vx;
}

/* ------------------------------------------------------------------ */
/* get_emf, FP64 (src/emfields.c:32-60)                                */
/* ------------------------------------------------------------------ */
template <int FK>
__device__ __forceinline__ void get_emf64(const FieldDev &f, double x, double y,
                                          double z, double t, Emf &o)
{
    if (FK == FK_TEST) {
        o.bx = 0.0; o.by = 0.0; o.bz = 1.0;
        o.ex = 0.0; o.ey = 0.0; o.ez = 0.0;
    } else if (FK == FK_ABC || FK == FK_ABCV) {
        /* getb_ff, src/force_free.c:138-150 */
        double sx, cx, sy, cy, sz, cz;
        tp_sincos(f.lam * x, &sx, &cx);
        tp_sincos(f.lam * y, &sy, &cy);
        tp_sincos(f.lam * z, &sz, &cz);
        double bx = f.A * sz + f.C * cy;
        double by = f.B * sx + f.A * cz;
        double bz = f.C * sy + f.B * cx;
        bx *= f.B0; by *= f.B0; bz *= f.B0;
        o.bx = bx; o.by = by; o.bz = bz;
        if (FK == FK_ABCV) {
            /* gete_ff, src/force_free.c:78-86, with the double velocity
             * gather of src/velocity_field.c:555-584 */
            unsigned c[8];
            double w[8];
            trilinear_ref(f.vnx, f.vny, f.vnz, f.vlo, f.vhi, f.vd, x, y, z, c, w);
            double vx = 0.0, vy = 0.0, vz = 0.0;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const double *v = f.vgrid + 3 * c[k];
                vx = vx + v[0] * w[k];
                vy = vy + v[1] * w[k];
                vz = vz + v[2] * w[k];
            }
            o.ex = by * vz - bz * vy;
            o.ey = bz * vx - bx * vz;
            o.ez = bx * vy - by * vx;
        } else {
            o.ex = 0.0; o.ey = 0.0; o.ez = 0.0;
        }
    } else if (FK == FK_GRID_REF) {
        /* getemf_pic_test_particle, src/emfields.c:153-166: float
         * accumulators rounded after every corner */
        unsigned c[8];
        double w[8];
        trilinear_ref(f.nx, f.ny, f.nz, f.lo, f.hi, f.d, x, y, z, c, w);
        float ax = 0.f, ay = 0.f, az = 0.f, bx = 0.f, by = 0.f, bz = 0.f;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            float4 p = f.grid[2 * c[k]];
            float4 q = f.grid[2 * c[k] + 1];
            ax = (float)((double)ax + (double)p.x * w[k]);
            ay = (float)((double)ay + (double)p.y * w[k]);
            az = (float)((double)az + (double)p.z * w[k]);
            bx = (float)((double)bx + (double)p.w * w[k]);
            by = (float)((double)by + (double)q.x * w[k]);
            bz = (float)((double)bz + (double)q.y * w[k]);
        }
        o.ex = ax; o.ey = ay; o.ez = az;
        o.bx = bx; o.by = by; o.bz = bz;
    } else if (FK == FK_GRID_TB) {
        unsigned c[8];
        double w[8];
        trilinear_tb(f.nx, f.ny, f.nz, f.lo, f.hi, f.d, x, y, z, c, w);
        double ax = 0.0, ay = 0.0, az = 0.0, bx = 0.0, by = 0.0, bz = 0.0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            float4 p = f.grid[2 * c[k]];
            float4 q = f.grid[2 * c[k] + 1];
            ax = ax + (double)p.x * w[k];
            ay = ay + (double)p.y * w[k];
            az = az + (double)p.z * w[k];
            bx = bx + (double)p.w * w[k];
            by = by + (double)q.x * w[k];
            bz = bz + (double)q.y * w[k];
        }
        o.ex = ax; o.ey = ay; o.ez = az;
        o.bx = bx; o.by = by; o.bz = bz;
    } else if (FK == FK_GRID2_REF || FK == FK_GRID2_TB) {
        /* two time slices (the reference's multi_tframe = 1,
         * src/electric_field.c:407-437, src/magnetic_field.c:593-622): each
         * slice summed over the corners in double, blended
         * dt * next + (1 - dt) * previous with dt the fraction of t / gdt
         * (src/interpolation.c:97-100); the reference gather returns float */
        unsigned c[8];
        double w[8];
        if (FK == FK_GRID2_REF) trilinear_ref(f.nx, f.ny, f.nz, f.lo, f.hi, f.d, x, y, z, c, w);
        else trilinear_tb(f.nx, f.ny, f.nz, f.lo, f.hi, f.d, x, y, z, c, w);
        double a[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, b[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float4 p = f.grid[2 * c[k]], q = f.grid[2 * c[k] + 1];
            const float4 pb = f.grid_b[2 * c[k]], qb = f.grid_b[2 * c[k] + 1];
            a[0] = a[0] + (double)p.x * w[k]; a[1] = a[1] + (double)p.y * w[k];
            a[2] = a[2] + (double)p.z * w[k]; a[3] = a[3] + (double)p.w * w[k];
            a[4] = a[4] + (double)q.x * w[k]; a[5] = a[5] + (double)q.y * w[k];
            b[0] = b[0] + (double)pb.x * w[k]; b[1] = b[1] + (double)pb.y * w[k];
            b[2] = b[2] + (double)pb.z * w[k]; b[3] = b[3] + (double)pb.w * w[k];
            b[4] = b[4] + (double)qb.x * w[k]; b[5] = b[5] + (double)qb.y * w[k];
        }
        const double r = t / f.gdt;
        const double dtf = r - (double)tp_floor_to_int(floor(r));
        double v[6];
#pragma unroll
        for (int k = 0; k < 6; k++) {
            v[k] = dtf * a[k] + (1.0 - dtf) * b[k];
            if (FK == FK_GRID2_REF) v[k] = (double)(float)v[k];
        }
        o.ex = v[0]; o.ey = v[1]; o.ez = v[2];
        o.bx = v[3]; o.by = v[4]; o.bz = v[5];
    } else if (FK == FK_MODES || FK == FK_MLP) {
        /* four independent cosines per iteration (instruction-level
         * parallelism for the FP64 FMA chains), summed in mode order */
        double bx = 0.0, by = 0.0, bz = 0.0;
        const int n4 = f.nmodes & ~3;
        /* the table is read-only for the whole kernel: read it through the
         * constant address space so uniform reads become scalar loads */
        typedef const double __attribute__((address_space(4))) *ctab;
        const ctab tab = (ctab)f.modes;
        for (int n = 0; n < n4; n += 4) {
            const ctab m = tab + 7 * n;
            double c0 = tp_cos_pi(fma(m[2], z, fma(m[1], y, fma(m[0], x, m[6]))));
            double c1 = tp_cos_pi(fma(m[9], z, fma(m[8], y, fma(m[7], x, m[13]))));
            double c2 = tp_cos_pi(fma(m[16], z, fma(m[15], y, fma(m[14], x, m[20]))));
            double c3 = tp_cos_pi(fma(m[23], z, fma(m[22], y, fma(m[21], x, m[27]))));
            bx = fma(m[3], c0, bx); by = fma(m[4], c0, by); bz = fma(m[5], c0, bz);
            bx = fma(m[10], c1, bx); by = fma(m[11], c1, by); bz = fma(m[12], c1, bz);
            bx = fma(m[17], c2, bx); by = fma(m[18], c2, by); bz = fma(m[19], c2, bz);
            bx = fma(m[24], c3, bx); by = fma(m[25], c3, by); bz = fma(m[26], c3, bz);
        }
        for (int n = n4; n < f.nmodes; n++) {
            const ctab m = tab + 7 * n;
            double cs = tp_cos_pi(fma(m[2], z, fma(m[1], y, fma(m[0], x, m[6]))));
            bx = fma(m[3], cs, bx);
            by = fma(m[4], cs, by);
            bz = fma(m[5], cs, bz);
        }
        o.bx = bx; o.by = by; o.bz = f.b0z + bz;
        o.ex = 0.0; o.ey = 0.0; o.ez = 0.0;
    } else if (FK == FK_WLCS) {
        emf_wlcs(f, x, y, z, t, o);
    } else if (FK == FK_MHD_REF) {
        get_emf64<FK_GRID_REF>(f, x, y, z, t, o);
        mhd_e64(o);
    } else if (FK == FK_GRIDD_REF) {
        /* get_double_efield_at_point / get_double_bfield_at_point
         * (src/electric_field.c:370-406, src/magnetic_field.c:555-591):
         * the reference's weights, double accumulators */
        unsigned c[8];
        double w[8];
        trilinear_ref(f.nx, f.ny, f.nz, f.lo, f.hi, f.d, x, y, z, c, w);
        double ax = 0.0, ay = 0.0, az = 0.0, bx = 0.0, by = 0.0, bz = 0.0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const double2 *nd = (const double2 *)(f.gridd + 8 * (size_t)c[k]);
            const double2 p = nd[0], q = nd[1], r = nd[2];
            ax = ax + p.x * w[k];
            ay = ay + p.y * w[k];
            az = az + q.x * w[k];
            bx = bx + q.y * w[k];
            by = by + r.x * w[k];
            bz = bz + r.y * w[k];
        }
        o.ex = ax; o.ey = ay; o.ez = az;
        o.bx = bx; o.by = by; o.bz = bz;
    }
}

/* FP32 sum over the modes: the phase is formed in revolutions from the
 * table's (k, phi) / 2 pi, reduced to [0, 1) by v_fract_f32 and passed to
 * the hardware cosine, whose argument is in revolutions (v_cos_f32:
 * cos(2 pi r)).  Four
 * modes per iteration for instruction-level parallelism; the table is
 * uniform and read through the constant address space (scalar loads). */
__device__ __forceinline__ void modes_sum32(const float *tab32, int nmodes, float x, float y,
                                            float z, float &bxo, float &byo, float &bzo)
{
    typedef const float __attribute__((address_space(4))) *ctab;
    const ctab t = (ctab)tab32;
    float bx = 0.f, by = 0.f, bz = 0.f;
    const int n4 = nmodes & ~3;
    for (int n = 0; n < n4; n += 4) {
        const ctab m = t + 8 * n;
        float u0 = fmaf(m[2], z, fmaf(m[1], y, fmaf(m[0], x, m[3])));
        float u1 = fmaf(m[10], z, fmaf(m[9], y, fmaf(m[8], x, m[11])));
        float u2 = fmaf(m[18], z, fmaf(m[17], y, fmaf(m[16], x, m[19])));
        float u3 = fmaf(m[26], z, fmaf(m[25], y, fmaf(m[24], x, m[27])));
        float c0 = __builtin_amdgcn_cosf(__builtin_amdgcn_fractf(u0));
        float c1 = __builtin_amdgcn_cosf(__builtin_amdgcn_fractf(u1));
        float c2 = __builtin_amdgcn_cosf(__builtin_amdgcn_fractf(u2));
        float c3 = __builtin_amdgcn_cosf(__builtin_amdgcn_fractf(u3));
        bx = fmaf(m[4], c0, bx); by = fmaf(m[5], c0, by); bz = fmaf(m[6], c0, bz);
        bx = fmaf(m[12], c1, bx); by = fmaf(m[13], c1, by); bz = fmaf(m[14], c1, bz);
        bx = fmaf(m[20], c2, bx); by = fmaf(m[21], c2, by); bz = fmaf(m[22], c2, bz);
        bx = fmaf(m[28], c3, bx); by = fmaf(m[29], c3, by); bz = fmaf(m[30], c3, bz);
    }
    for (int n = n4; n < nmodes; n++) {
        const ctab m = t + 8 * n;
        float u = fmaf(m[2], z, fmaf(m[1], y, fmaf(m[0], x, m[3])));
        float c = __builtin_amdgcn_cosf(__builtin_amdgcn_fractf(u));
        bx = fmaf(m[4], c, bx); by = fmaf(m[5], c, by); bz = fmaf(m[6], c, bz);
    }
    bxo = bx; byo = by; bzo = bz;
}

/* ------------------------------------------------------------------ */
/* get_emf, FP32 (precision form)                                      */
/* ------------------------------------------------------------------ */
/* The FP64 reference gather at one point (bit for bit get_emf64<FK>) for a
 * kernel whose cache was filled by the FP32 form's cell selection: the cell
 * is selected again with the FP64 form's divisions, and when the cache holds
 * that cell its floats are summed; otherwise the nodes are read. */
template <int FK, int NS>
__device__ __forceinline__ void ref_emf64_via_cache(const FieldDev &f, const GridCacheN<NS> &gc,
                                                    double x, double y, double z, double t,
                                                    Emf &o)
{
    int i1[3], i2[3];
    double fx, fy, fz;
    ref_cell_index<false>(f, x, y, z, i1, i2, fx, fy, fz);
    const bool same = gc.valid && i1[0] == gc.key[0] && i2[0] == gc.key[1] &&
                      i1[1] == gc.key[2] && i2[1] == gc.key[3] && i1[2] == gc.key[4] &&
                      i2[2] == gc.key[5];
    if (same) {
        double w[8];
        ref_weights<double>(fx, fy, fz, w);
        if constexpr (NS == 1) ref_sum64(gc, w, o);
        else ref_sum64(gc, w, f.gdt, t, true, o);
    } else {
        get_emf64<FK>(f, x, y, z, t, o);
    }
}

template <int FK>
__device__ __forceinline__ void get_emf32(const FieldDev &f, double x, double y,
                                          double z, double t, EmfF &o)
{
    if (FK == FK_TEST) {
        o.bx = 0.f; o.by = 0.f; o.bz = 1.f;
        o.ex = 0.f; o.ey = 0.f; o.ez = 0.f;
    } else if (FK == FK_ABC) {
        /* the hardware sine and cosine of the phase in revolutions (fraction
         * taken first), as the FP32 mode sum evaluates its cosines */
        const float lr = (float)(f.lam * (0.5 / TP_PI));
        const float px = __builtin_amdgcn_fractf(lr * (float)x);
        const float py = __builtin_amdgcn_fractf(lr * (float)y);
        const float pz = __builtin_amdgcn_fractf(lr * (float)z);
        const float sx = __builtin_amdgcn_sinf(px), cx = __builtin_amdgcn_cosf(px);
        const float sy = __builtin_amdgcn_sinf(py), cy = __builtin_amdgcn_cosf(py);
        const float sz = __builtin_amdgcn_sinf(pz), cz = __builtin_amdgcn_cosf(pz);
        float A = (float)f.A, B = (float)f.B, C = (float)f.C, B0 = (float)f.B0;
        o.bx = (A * sz + C * cy) * B0;
        o.by = (B * sx + A * cz) * B0;
        o.bz = (C * sy + B * cx) * B0;
        o.ex = 0.f; o.ey = 0.f; o.ez = 0.f;
    } else if (FK == FK_GRID_TB || FK == FK_GRID_REF) {
        /* float weights and accumulation; the weight convention follows
         * the field's gather mode */
        unsigned c[8];
        double w[8];
        if (FK == FK_GRID_REF)
            trilinear_ref(f.nx, f.ny, f.nz, f.lo, f.hi, f.d, x, y, z, c, w);
        else
            trilinear_tb(f.nx, f.ny, f.nz, f.lo, f.hi, f.d, x, y, z, c, w);
        float ax = 0.f, ay = 0.f, az = 0.f, bx = 0.f, by = 0.f, bz = 0.f;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            float4 p = f.grid[2 * c[k]];
            float4 q = f.grid[2 * c[k] + 1];
            float wk = (float)w[k];
            ax = fmaf(p.x, wk, ax); ay = fmaf(p.y, wk, ay); az = fmaf(p.z, wk, az);
            bx = fmaf(p.w, wk, bx); by = fmaf(q.x, wk, by); bz = fmaf(q.y, wk, bz);
        }
        o.ex = ax; o.ey = ay; o.ez = az;
        o.bx = bx; o.by = by; o.bz = bz;
    } else if (FK == FK_GRID2_TB || FK == FK_GRID2_REF) {
        /* both slices in float, blended in float */
        unsigned c[8];
        double w[8];
        if (FK == FK_GRID2_REF)
            trilinear_ref(f.nx, f.ny, f.nz, f.lo, f.hi, f.d, x, y, z, c, w);
        else
            trilinear_tb(f.nx, f.ny, f.nz, f.lo, f.hi, f.d, x, y, z, c, w);
        float a[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, b[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float wk = (float)w[k];
            const float4 p = f.grid[2 * c[k]], q = f.grid[2 * c[k] + 1];
            const float4 pb = f.grid_b[2 * c[k]], qb = f.grid_b[2 * c[k] + 1];
            a[0] = fmaf(p.x, wk, a[0]); a[1] = fmaf(p.y, wk, a[1]); a[2] = fmaf(p.z, wk, a[2]);
            a[3] = fmaf(p.w, wk, a[3]); a[4] = fmaf(q.x, wk, a[4]); a[5] = fmaf(q.y, wk, a[5]);
            b[0] = fmaf(pb.x, wk, b[0]); b[1] = fmaf(pb.y, wk, b[1]); b[2] = fmaf(pb.z, wk, b[2]);
            b[3] = fmaf(pb.w, wk, b[3]); b[4] = fmaf(qb.x, wk, b[4]); b[5] = fmaf(qb.y, wk, b[5]);
        }
        const double r = t / f.gdt;
        const float dtf = (float)(r - floor(r));
        const float dtg = 1.0f - dtf;
        o.ex = dtf * a[0] + dtg * b[0]; o.ey = dtf * a[1] + dtg * b[1];
        o.ez = dtf * a[2] + dtg * b[2]; o.bx = dtf * a[3] + dtg * b[3];
        o.by = dtf * a[4] + dtg * b[4]; o.bz = dtf * a[5] + dtg * b[5];
    } else if (FK == FK_MODES || FK == FK_MLP) {
        float bx, by, bz;
        modes_sum32(f.modes32, f.nmodes, (float)x, (float)y, (float)z, bx, by, bz);
        o.bx = bx; o.by = by; o.bz = (float)f.b0z + bz;
        o.ex = 0.f; o.ey = 0.f; o.ez = 0.f;
    } else if (FK == FK_MHD_REF) {
        get_emf32<FK_GRID_REF>(f, x, y, z, t, o);
        mhd_e32(o);
    } else {
        Emf e;
        get_emf64<FK>(f, x, y, z, t, e);
        o.bx = (float)e.bx; o.by = (float)e.by; o.bz = (float)e.bz;
        o.ex = (float)e.ex; o.ey = (float)e.ey; o.ez = (float)e.ez;
    }
}

/* field evaluation inside the step loop: grids go through the cell cache */
template <int FK, class GC>
__device__ __forceinline__ void emf64_step(const FieldDev &f, GC &gc, double x,
                                           double y, double z, double t, Emf &o)
{
    if constexpr (FK == FK_GRID_REF || FK == FK_GRID_TB) {
        grid_emf64_cached<FK>(f, gc, x, y, z, o);
    } else if constexpr (FK == FK_GRID2_REF || FK == FK_GRID2_TB) {
        grid2_emf64_cached<FK>(f, gc, x, y, z, t, o);
    } else if constexpr (FK == FK_MHD_REF) {
        grid_emf64_cached<FK_GRID_REF>(f, gc, x, y, z, o);
        mhd_e64(o);
    } else {
        get_emf64<FK>(f, x, y, z, t, o);
    }
}

template <int FK, class GC>
__device__ __forceinline__ void emf32_step(const FieldDev &f, GC &gc, double x,
                                           double y, double z, double t, EmfF &o)
{
    if constexpr (FK == FK_GRID_REF || FK == FK_GRID_TB) {
        grid_emf32_cached<FK>(f, gc, x, y, z, o);
    } else if constexpr (FK == FK_GRID2_REF || FK == FK_GRID2_TB) {
        grid2_emf32_cached<FK>(f, gc, x, y, z, t, o);
    } else if constexpr (FK == FK_MHD_REF) {
        grid_emf32_cached<FK_GRID_REF>(f, gc, x, y, z, o);
        mhd_e32(o);
    } else {
        get_emf32<FK>(f, x, y, z, t, o);
    }
}

/* ------------------------------------------------------------------ */
/* pushers                                                             */
/* ------------------------------------------------------------------ */
struct Ptl {
    double x, y, z, ux, uy, uz, t;
};

/* tracking_wirz, src/tracking.c:749-892, in the reference's order */
template <int FK, class GC>
__device__ __forceinline__ void wirz_step64(const FieldDev &f, GC &gc, Ptl &p,
                                            double dt, double qm)
{
    Emf e;
    emf64_step<FK>(f, gc, p.x, p.y, p.z, p.t, e);
    double Btot = sqrt(e.bx * e.bx + e.by * e.by + e.bz * e.bz);
    double iB = 1.0 / Btot;
    double dth = dt * 0.5;
    double qd = qm * dth;
    double dmx = p.ux + qd * e.ex;
    double dmy = p.uy + qd * e.ey;
    double dmz = p.uz + qd * e.ez;
    double gama = sqrt(1.0 + dmx * dmx + dmy * dmy + dmz * dmz);
    double ig = 1.0 / gama;
    double vmx = dmx * ig, vmy = dmy * ig, vmz = dmz * ig;
    double hx = e.bx * iB, hy = e.by * iB, hz = e.bz * iB;
    double vd = 0.0;
    vd += vmx * hx; vd += vmy * hy; vd += vmz * hz;
    double ax = vmx - vd * hx, ay = vmy - vd * hy, az = vmz - vd * hz;
    double vp = sqrt(ax * ax + ay * ay + az * az);
    double px = (vmx - vd * hx) / vp, py = (vmy - vd * hy) / vp, pz = (vmz - vd * hz) / vp;
    double rx = hy * pz - hz * py, ry = hz * px - hx * pz, rz = hx * py - hy * px;
    double gR = vp * gama * 1.0 * iB / qm;
    double om = qm * Btot * ig;
    double dthe = om * dt;
    double sd, cd;
    tp_sincos(dthe * 0.5, &sd, &cd);
    double t1 = dth * vd * 1.0;
    double t2 = gR * sd;
    double t3 = -gR * (1.0 - cd);
    double prx = p.x + hx * t1 + px * t2 + rx * t3;
    double pry = p.y + hy * t1 + py * t2 + ry * t3;
    double prz = p.z + hz * t1 + pz * t2 + rz * t3;

    emf64_step<FK>(f, gc, prx, pry, prz, p.t + dth, e);
    Btot = sqrt(e.bx * e.bx + e.by * e.by + e.bz * e.bz);
    iB = 1.0 / Btot;
    hx = e.bx * iB; hy = e.by * iB; hz = e.bz * iB;
    vd = 0.0;
    vd += vmx * hx; vd += vmy * hy; vd += vmz * hz;
    ax = vmx - vd * hx; ay = vmy - vd * hy; az = vmz - vd * hz;
    vp = sqrt(ax * ax + ay * ay + az * az);
    px = (vmx - vd * hx) / vp; py = (vmy - vd * hy) / vp; pz = (vmz - vd * hz) / vp;
    rx = hy * pz - hz * py; ry = hz * px - hx * pz; rz = hx * py - hy * px;
    om = qm * Btot * ig;
    dthe = om * dt;
    tp_sincos(dthe, &sd, &cd);
    double wx = hx * vd + vp * (cd * px - sd * rx);
    double wy = hy * vd + vp * (cd * py - sd * ry);
    double wz = hz * vd + vp * (cd * pz - sd * rz);
    /* get_delta_vec, src/tracking.c:1066-1078 */
    double bm = 0.0;
    bm += wx * wx; bm += wy * wy; bm += wz * wz;
    double g2 = 1.0 / sqrt(1.0 - bm);
    p.ux = g2 * wx + qd * e.ex;
    p.uy = g2 * wy + qd * e.ey;
    p.uz = g2 * wz + qd * e.ez;
    vd = 0.0;
    vd += wx * hx; vd += wy * hy; vd += wz * hz;
    ax = wx - vd * hx; ay = wy - vd * hy; az = wz - vd * hz;
    vp = sqrt(ax * ax + ay * ay + az * az);
    gR = vp * gama * 1.0 * iB / qm;
    t1 = dt * vd * 1.0;
    t2 = gR * sd;
    t3 = -gR * (1.0 - cd);
    p.x += hx * t1 + px * t2 + rx * t3;
    p.y += hy * t1 + py * t2 + ry * t3;
    p.z += hz * t1 + pz * t2 + rz * t3;
    p.t += dt;
}

/* relativistic Boris (same operations as the oracle's tpref_boris_step) */
template <int FK, class GC>
__device__ __forceinline__ void boris_step64(const FieldDev &f, GC &gc, Ptl &p,
                                             double dt, double qm)
{
    Emf e;
    emf64_step<FK>(f, gc, p.x, p.y, p.z, p.t, e);
    double h = qm * (0.5 * dt);
    double umx = p.ux + h * e.ex;
    double umy = p.uy + h * e.ey;
    double umz = p.uz + h * e.ez;
    double gm = sqrt(1.0 + umx * umx + umy * umy + umz * umz);
    double fq = h / gm;
    double tx = fq * e.bx, ty = fq * e.by, tz = fq * e.bz;
    double t2 = tx * tx + ty * ty + tz * tz;
    double sf = 2.0 / (1.0 + t2);
    double sx = sf * tx, sy = sf * ty, sz = sf * tz;
    double upx = umx + (umy * tz - umz * ty);
    double upy = umy + (umz * tx - umx * tz);
    double upz = umz + (umx * ty - umy * tx);
    double uqx = umx + (upy * sz - upz * sy);
    double uqy = umy + (upz * sx - upx * sz);
    double uqz = umz + (upx * sy - upy * sx);
    double ux = uqx + h * e.ex;
    double uy = uqy + h * e.ey;
    double uz = uqz + h * e.ez;
    double g = sqrt(1.0 + ux * ux + uy * uy + uz * uz);
    double ig = dt / g;
    p.ux = ux; p.uy = uy; p.uz = uz;
    p.x += ux * ig;
    p.y += uy * ig;
    p.z += uz * ig;
    p.t += dt;
}

/* Boris in FP32 with FP64 position and time accumulators, given the field
 * at the particle */
__device__ __forceinline__ void boris_push32(const EmfF &e, Ptl &p, float &ux, float &uy,
                                             float &uz, double dt, float dtf, float hf)
{
    float umx = fmaf(hf, e.ex, ux);
    float umy = fmaf(hf, e.ey, uy);
    float umz = fmaf(hf, e.ez, uz);
    float gm2 = fmaf(umx, umx, fmaf(umy, umy, fmaf(umz, umz, 1.0f)));
    float fq = hf * __frsqrt_rn(gm2);
    float tx = fq * e.bx, ty = fq * e.by, tz = fq * e.bz;
    float t2 = fmaf(tx, tx, fmaf(ty, ty, tz * tz));
    float sf = 2.0f * __builtin_amdgcn_rcpf(1.0f + t2);   /* hardware reciprocal, ~1 ulp */
    float sx = sf * tx, sy = sf * ty, sz = sf * tz;
    float upx = umx + (umy * tz - umz * ty);
    float upy = umy + (umz * tx - umx * tz);
    float upz = umz + (umx * ty - umy * tx);
    float uqx = umx + (upy * sz - upz * sy);
    float uqy = umy + (upz * sx - upx * sz);
    float uqz = umz + (upx * sy - upy * sx);
    ux = fmaf(hf, e.ex, uqx);
    uy = fmaf(hf, e.ey, uqy);
    uz = fmaf(hf, e.ez, uqz);
    float g2 = fmaf(ux, ux, fmaf(uy, uy, fmaf(uz, uz, 1.0f)));
    float ig = dtf * __frsqrt_rn(g2);
    p.x += (double)(ux * ig);
    p.y += (double)(uy * ig);
    p.z += (double)(uz * ig);
    p.t += dt;
}

template <int FK, class GC>
__device__ __forceinline__ void boris_step32(const FieldDev &f, GC &gc, Ptl &p,
                                             float &ux, float &uy, float &uz,
                                             double dt, float dtf, float hf)
{
    EmfF e;
    emf32_step<FK>(f, gc, p.x, p.y, p.z, p.t, e);
    boris_push32(e, p, ux, uy, uz, dt, dtf, hf);
}

#endif /* TP_DEVICE_H */
