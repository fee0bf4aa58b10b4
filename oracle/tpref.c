/*
 * tpref.c -- the project's FP64 CPU oracle.  TEST INFRASTRUCTURE ONLY:
 * imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg as the checker; never part of the product path.
 *
 * A restatement, in plain C, of the reference tracer's hot path
 * (xiaocanli/test_particle):
 *   fields     get_emf src/emfields.c:32-60, getemf_test :210-218,
 *              getb_ff/gete_ff src/force_free.c:78-86, :138-150,
 *              trilinear gather src/interpolation.c:36-181 with the float
 *              gathers src/magnetic_field.c:593-622,
 *              src/electric_field.c:408-437 and the double velocity gather
 *              src/velocity_field.c:555-584,
 *              WLCS src/wlcs.c:281-449, :814-903
 *   pusher     tracking_wirz src/tracking.c:749-892 (+ helpers :1066-1140)
 *   bc         particle_bc src/tracking.c:1172-1221
 *   loop       src/tracking.c:90-190, :297-462
 *   diagnostics ptl_energy_fixed src/diagnostics.c:187-205,
 *              calc_diff_coeff :659-705
 *   poincare   the Poincare map (run_type 2): tracing loop
 *              src/poincare_map.c:83-185, next_location :268-320,
 *              field_boundary :326-352
 * plus, for what the reference does not have, a relativistic Boris push
 * and a multi-mode turbulence field.
 *
 * Arithmetic follows the reference's operation order, without
 * contraction (compiled -ffp-contract=off), so that with libm sin/cos
 * (tpref_set_math(1)) it reproduces the reference binary bit for bit, and
 * with the shared tp_sincos (tpref_set_math(0)) the GPU kernels.
 *
 * Pins (tests/): golden vectors written by the reference binary through
 * oracle/ref_harness.c (fields of systems 0/2/4, Wirz trajectories,
 * calc_diff_coeff, spectrum/D sums of the full loop, Poincare crossing
 * points); closed forms (Wirz is
 * exact in uniform B; Boris rotation angle 2 atan(Omega dt / 2 gamma));
 * scipy ellipk/ellipe for the elliptic integrals; brute-force mode sums.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "tpref.h"
#include "tp_math.h"

static int g_use_libm = 1;
static int g_carry_stale = 0;

void tpref_set_math(int use_libm) { g_use_libm = use_libm; }

/* 1: the stale per-particle D terms (dxx dyy dzz daa dee) carry over from
 * one particle to the next in loop order, as the reference's loop-private
 * variables do with one thread (src/tracking.c:306-312, 334-336); 0: each
 * particle starts from zero (the semantics the GPU kernels implement). */
void tpref_set_carry_stale(int on) { g_carry_stale = on; }

void tpref_sincos(double x, double *s, double *c) { tp_sincos(x, s, c); }

static void ellint_ref(double k, double *elle, double *ellk);
void tpref_ellint_ref(double k, double *elle, double *ellk) { ellint_ref(k, elle, ellk); }

static inline double o_sin(double x) { return g_use_libm ? sin(x) : tp_sin(x); }
static inline double o_cos(double x) { return g_use_libm ? cos(x) : tp_cos(x); }

typedef struct { double Bx, By, Bz, Ex, Ey, Ez; } emf_t;

/* ---------------------------------------------------------------------- */
/* trilinear gather                                                        */
/* ---------------------------------------------------------------------- */

/* grid_indices + get_trilinar_parameters, src/interpolation.c:36-105,
 * 141-181.  Corner order and weights as there (the x and z fractions are
 * exchanged relative to the corner they belong to).  The flat corner
 * index is clamped into the array; the reference reads past a row when a
 * coordinate equals the upper bound exactly. */
static void trilinear_ref(int nx, int ny, int nz, const double lo[3],
                          const double hi[3], double x, double y, double z,
                          long corners[8], double w[8])
{
    double ddx = (hi[0] - lo[0]) / (nx - 1.0);
    double ddy = (hi[1] - lo[1]) / (ny - 1.0);
    double ddz = (hi[2] - lo[2]) / (nz - 1.0);
    double fx, fy, fz, vx, vy, vz;
    int ix1, iy1, iz1, ix2, iy2, iz2;
    vx = floor((x - lo[0]) / ddx);
    vy = floor((y - lo[1]) / ddy);
    vz = floor((z - lo[2]) / ddz);
    /* (int) of a value outside int range is INT_MIN on x86-64 */
    ix1 = (vx > -2147483649.0 && vx < 2147483648.0) ? (int)vx : (-2147483647 - 1);
    iy1 = (vy > -2147483649.0 && vy < 2147483648.0) ? (int)vy : (-2147483647 - 1);
    iz1 = (vz > -2147483649.0 && vz < 2147483648.0) ? (int)vz : (-2147483647 - 1);
    ix2 = ix1 + 1;
    iy2 = iy1 + 1;
    iz2 = iz1 + 1;
    fx = (x - lo[0]) / ddx - ix1;
    fy = (y - lo[1]) / ddy - iy1;
    fz = (z - lo[2]) / ddz - iz1;
    if (x > hi[0]) { ix1 = nx - 1; ix2 = 0; fx = x / ddx - (nx - 1); }
    if (y > hi[1]) { iy1 = ny - 1; iy2 = 0; fy = y / ddy - (ny - 1); }
    if (z > hi[2]) { iz1 = nz - 1; iz2 = 0; fz = z / ddz - (nz - 1); }
    if (x < lo[0]) { ix1 = nx - 1; ix2 = 0; fx = x / ddx + 1.0; }
    if (y < lo[1]) { iy1 = ny - 1; iy2 = 0; fy = y / ddy + 1.0; }
    if (z < lo[2]) { iz1 = nz - 1; iz2 = 0; fz = z / ddz + 1.0; }
    long total = (long)nx * ny * nz;
#define FLAT(a, b, c) ((long)(c) + (long)nx * ((long)(b) + (long)ny * (long)(a)))
    long cc[8] = {
        FLAT(iz1, iy1, ix1), FLAT(iz2, iy1, ix1), FLAT(iz1, iy2, ix1),
        FLAT(iz1, iy1, ix2), FLAT(iz1, iy2, ix2), FLAT(iz2, iy1, ix2),
        FLAT(iz2, iy2, ix1), FLAT(iz2, iy2, ix2)};
#undef FLAT
    for (int i = 0; i < 8; i++) {
        long c = cc[i];
        corners[i] = c < 0 ? 0 : (c >= total ? total - 1 : c);
    }
    w[0] = (1.0 - fx) * (1.0 - fy) * (1.0 - fz);
    w[1] = fx * (1.0 - fy) * (1.0 - fz);
    w[2] = (1.0 - fx) * fy * (1.0 - fz);
    w[3] = (1.0 - fx) * (1.0 - fy) * fz;
    w[4] = (1.0 - fx) * fy * fz;
    w[5] = fx * (1.0 - fy) * fz;
    w[6] = fx * fy * (1.0 - fz);
    w[7] = fx * fy * fz;
}

/* Textbook trilinear on the same endpoint-inclusive nodes: the position is
 * folded into [lo, hi) by the period hi - lo, corner ix+1 gets weight fx. */
static void trilinear_textbook(int nx, int ny, int nz, const double lo[3],
                               const double hi[3], double x, double y,
                               double z, long corners[8], double w[8])
{
    double p[3] = {x, y, z};
    int n[3] = {nx, ny, nz};
    int i1[3], i2[3];
    double f[3];
    for (int a = 0; a < 3; a++) {
        double L = hi[a] - lo[a];
        double d = L / (n[a] - 1.0);
        double q = p[a] - lo[a];
        q = q - L * floor(q / L);
        double v = floor(q / d);
        int i = (v >= 0.0 && v <= n[a] - 2.0) ? (int)v : (v < 0.0 ? 0 : n[a] - 2);
        if (!(v == v)) i = 0;
        i1[a] = i;
        i2[a] = i + 1;
        f[a] = q / d - i;
    }
    /* same corner order as the reference */
#define FLAT(a, b, c) ((long)(c) + (long)nx * ((long)(b) + (long)ny * (long)(a)))
    corners[0] = FLAT(i1[2], i1[1], i1[0]);
    corners[1] = FLAT(i2[2], i1[1], i1[0]);
    corners[2] = FLAT(i1[2], i2[1], i1[0]);
    corners[3] = FLAT(i1[2], i1[1], i2[0]);
    corners[4] = FLAT(i1[2], i2[1], i2[0]);
    corners[5] = FLAT(i2[2], i1[1], i2[0]);
    corners[6] = FLAT(i2[2], i2[1], i1[0]);
    corners[7] = FLAT(i2[2], i2[1], i2[0]);
#undef FLAT
    double fx = f[0], fy = f[1], fz = f[2];
    w[0] = (1.0 - fx) * (1.0 - fy) * (1.0 - fz);
    w[1] = (1.0 - fx) * (1.0 - fy) * fz;
    w[2] = (1.0 - fx) * fy * (1.0 - fz);
    w[3] = fx * (1.0 - fy) * (1.0 - fz);
    w[4] = fx * fy * (1.0 - fz);
    w[5] = fx * (1.0 - fy) * fz;
    w[6] = (1.0 - fx) * fy * fz;
    w[7] = fx * fy * fz;
}

/* get_float_{e,b}field_at_point: float accumulator, rounded after every
 * corner (src/magnetic_field.c:603-608) */
static void gather_float_ref(const float *g, const long c[8], const double w[8],
                             float out[3])
{
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    for (int i = 0; i < 8; i++) {
        ax = (float)((double)ax + (double)g[3 * c[i]] * w[i]);
        ay = (float)((double)ay + (double)g[3 * c[i] + 1] * w[i]);
        az = (float)((double)az + (double)g[3 * c[i] + 2] * w[i]);
    }
    out[0] = ax; out[1] = ay; out[2] = az;
}

static void gather_double(const float *gf, const double *gd, const long c[8],
                          const double w[8], double out[3])
{
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int i = 0; i < 8; i++) {
        double vx = gf ? (double)gf[3 * c[i]] : gd[3 * c[i]];
        double vy = gf ? (double)gf[3 * c[i] + 1] : gd[3 * c[i] + 1];
        double vz = gf ? (double)gf[3 * c[i] + 2] : gd[3 * c[i] + 2];
        ax = ax + vx * w[i];
        ay = ay + vy * w[i];
        az = az + vz * w[i];
    }
    out[0] = ax; out[1] = ay; out[2] = az;
}

/* ---------------------------------------------------------------------- */
/* WLCS                                                                    */
/* ---------------------------------------------------------------------- */
#define WL_L0 6.96E8     /* include/constants.h:14 */
#define WL_C0 3.0E10     /* include/constants.h:4 */
#define WL_I0 4.0E8      /* include/wlcs.h:41 */

/* Complete elliptic integrals K(k), E(k) by the arithmetic-geometric mean
 * (Abramowitz & Stegun 17.6.1-17.6.4).  The reference evaluates the same
 * functions with Fukushima's piecewise series (src/wlcs.c:460-809); both
 * are accurate to about 1e-15. */
void tpref_ellint(double k, double *elle, double *ellk)
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
    double K = M_PI / (2.0 * a);
    *ellk = K;
    *elle = K * (1.0 - sum);
}

/* ellint as the reference evaluates it.  For m = k^2 < 0.9 its Taylor
 * series are accurate and equal K, E.  For m >= 0.9 (src/wlcs.c:741-806) it
 * uses the nome relation K = -ln q(1-m) K'/pi and Legendre's relation
 * E = (pi/2 + K H')/K', but evaluates the K' series about 0.05 at offset
 * m - 0.95 instead of 0.95 - m, i.e. it takes K' = K(m - 0.9) where
 * K(1 - m) belongs.  Restated: K = K(m) K(m-0.9) / K(1-m),
 * E = (pi/2 + K (K(1-m) - E(1-m))) / K(m-0.9). */
static void ellint_ref(double k, double *elle, double *ellk)
{
    double m = k * k;
    if (!(m >= 0.9)) {
        tpref_ellint(k, elle, ellk);
        return;
    }
    double Ec, Kc, Em, Km, Ew, Kw;
    tpref_ellint(sqrt(1.0 - m), &Ec, &Kc);
    tpref_ellint(k, &Em, &Km);
    tpref_ellint(sqrt(m - 0.9), &Ew, &Kw);
    double Kref = (M_PI * Km / Kc) * (Kw / M_PI);
    *ellk = Kref;
    *elle = 1.0 / Kw * (M_PI / 2.0 + Kref * (Kc - Ec));
}

/* transform / reverse_tran, src/wlcs.c:871-903 */
static void wl_transform(double *ax, double *ay, double *az, double c1,
                         double s1, double c2, double s2)
{
    double ax1 = c2 * (*ax) + s2 * (*ay);
    double ay1 = -s2 * (*ax) + c2 * (*ay);
    double az1 = *az;
    *ax = c1 * ax1 - s1 * az1;
    *ay = ay1;
    *az = s1 * ax1 + c1 * az1;
}

static void wl_reverse(double *ax, double *ay, double *az, double c1,
                       double s1, double c2, double s2)
{
    double ax1 = c1 * (*ax) + s1 * (*az);
    double ay1 = *ay;
    double az1 = -s1 * (*ax) + c1 * (*az);
    *ax = c2 * ax1 - s2 * ay1;
    *ay = s2 * ax1 + c2 * ay1;
    *az = az1;
}

typedef struct {
    double cur_wr, x_wr, y_wr, z_wr, ct, st, cp, sp, t0;
    double cur_lp, r_lp, x_lp, y_lp, z_lp, ca, sa, cb, sb, omega;
} wl_sys;

/* the table read_wlcs builds (src/wlcs.c:166-212) */
static void wl_setup(const tpref_field *f, int i, wl_sys *s)
{
    const double *r = f->wlcs_rows + 14 * i;
    s->cur_wr = r[0] * f->norm_i;
    s->x_wr = r[1]; s->y_wr = r[2]; s->z_wr = r[3];
    s->ct = o_cos(r[4]); s->st = o_sin(r[4]);
    s->cp = o_cos(r[5]); s->sp = o_sin(r[5]);
    s->t0 = r[6];
    s->cur_lp = r[7] * f->norm_i;
    s->r_lp = r[8]; s->x_lp = r[9]; s->y_lp = r[10]; s->z_lp = r[11];
    s->ca = o_cos(r[12]); s->sa = o_sin(r[12]);
    s->cb = o_cos(r[13]); s->sb = o_sin(r[13]);
    s->omega = f->omega;
}

/* getWireB, src/wlcs.c:361-401 */
static void wl_wire_static(const wl_sys *s, double x, double y, double z,
                           double b[3])
{
    double xp = x - s->x_wr, yp = y - s->y_wr, zp = z - s->z_wr;
    wl_transform(&xp, &yp, &zp, s->ct, s->st, s->cp, s->sp);
    double rho = sqrt(xp * xp + yp * yp);
    if (rho < 0.1) rho = 0.1;
    double cphi = xp / rho;
    double phi0 = (yp >= 0.0) ? acos(cphi) : 2.0 * M_PI - acos(cphi);
    double bphi = 1.0 / 5.0 * (s->cur_wr * WL_I0) / (rho * WL_L0);
    double bx = bphi * o_cos(phi0 + M_PI / 2.0);
    double by = bphi * o_sin(phi0 + M_PI / 2.0);
    double bz = 0.0;
    wl_reverse(&bx, &by, &bz, s->ct, s->st, s->cp, s->sp);
    b[0] = bx; b[1] = by; b[2] = bz;
}

/* getWire_Bessel, src/wlcs.c:281-355 (float Bessel functions as there) */
static void wl_wire_bessel(const wl_sys *s, double x, double y, double z,
                           double t, double b[3], double e[3])
{
    double xp = x - s->x_wr, yp = y - s->y_wr, zp = z - s->z_wr;
    double t1 = s->t0 + t;
    double om = s->omega;
    wl_transform(&xp, &yp, &zp, s->ct, s->st, s->cp, s->sp);
    double rho = sqrt(xp * xp + yp * yp);
    double x0 = om * rho * WL_L0 / WL_C0;
    double bj0 = j0f(x0), bj1 = j1f(x0), by0 = y0f(x0), by1 = y1f(x0);
    double const0 = (s->cur_wr * WL_I0) / (5.0 * rho * WL_L0) * M_PI / 2.0;
    double wb = const0 * x0 * (-by1 * o_cos(om * t1) + bj1 * o_sin(om * t1));
    double phi1;
    if (rho == 0.0) phi1 = M_PI / 2.0;
    else if (yp >= 0.0) phi1 = acos(xp / rho) + M_PI / 2.0;
    else phi1 = 2.0 * M_PI - acos(xp / rho) + M_PI / 2.0;
    double bx = wb * o_cos(phi1), by = wb * o_sin(phi1), bz = 0.0;
    wl_reverse(&bx, &by, &bz, s->ct, s->st, s->cp, s->sp);
    double we = -const0 * x0 * (by0 * o_sin(om * t1) + bj0 * o_cos(om * t1));
    double ex = 0.0, ey = 0.0, ez = we;
    wl_reverse(&ex, &ey, &ez, s->ct, s->st, s->cp, s->sp);
    b[0] = bx; b[1] = by; b[2] = bz;
    e[0] = ex; e[1] = ey; e[2] = ez;
}

/* getLoopB, src/wlcs.c:406-449 */
static void wl_loop(const wl_sys *s, int compat, double x, double y, double z,
                    double b[3])
{
    double xp = (x - s->x_lp) * WL_L0;
    double yp = (y - s->y_lp) * WL_L0;
    double zp = (z - s->z_lp) * WL_L0;
    double al = s->r_lp * WL_L0;
    double curI = s->cur_lp * WL_I0;
    wl_transform(&xp, &yp, &zp, s->cb, s->sb, s->ca, s->sa);
    double Irho = 0.2 * curI;
    double rho_xy = sqrt(xp * xp + yp * yp);
    double al2 = al * al;
    double r2 = rho_xy * rho_xy + zp * zp;
    double irho1 = 1 / sqrt((al + rho_xy) * (al + rho_xy) + zp * zp);
    double irho2 = 1 / ((al - rho_xy) * (al - rho_xy) + zp * zp);
    double k = sqrt(4.0 * al * rho_xy) * irho1;
    double elle, ellk;
    if (compat) ellint_ref(k, &elle, &ellk);
    else tpref_ellint(k, &elle, &ellk);
    double Brho = Irho * zp * irho1 * (-ellk + (al2 + r2) * elle * irho2) / rho_xy;
    if (rho_xy == 0.0) Brho = 0.0;
    double Bz = Irho * irho1 * (ellk - (r2 - al2) * elle * irho2);
    Brho /= rho_xy;
    double bx = Brho * xp, by = Brho * yp, bz = Bz;
    wl_reverse(&bx, &by, &bz, s->cb, s->sb, s->ca, s->sa);
    b[0] = bx; b[1] = by; b[2] = bz;
}

/* getemf_wlcs, src/wlcs.c:814-865 */
static void emf_wlcs(const tpref_field *f, double x, double y, double z,
                     double t, emf_t *o)
{
    double Bx = 0.0, By = 0.0, Bz = 0.0, Ex = 0.0, Ey = 0.0, Ez = 0.0;
    for (int i = 0; i < f->nwlcs; i++) {
        wl_sys s;
        double b[3], e[3];
        wl_setup(f, i, &s);
        if (f->ibessel == 0) {
            wl_wire_static(&s, x, y, z, b);
            Bx += b[0]; By += b[1]; Bz += b[2];
        } else {
            wl_wire_bessel(&s, x, y, z, t, b, e);
            Bx += b[0]; Ex += e[0];
            By += b[1]; Ey += e[1];
            Bz += b[2]; Ez += e[2];
        }
        wl_loop(&s, f->wlcs_mode == TP_GATHER_REF_COMPAT, x, y, z, b);
        Bx += b[0]; By += b[1]; Bz += b[2];
    }
    o->Bx = Bx; o->By = By; o->Bz = Bz + 0.0;
    o->Ex = Ex; o->Ey = Ey; o->Ez = Ez;
}

/* ---------------------------------------------------------------------- */
/* get_emf                                                                 */
/* ---------------------------------------------------------------------- */
static void get_emf(const tpref_field *f, double x, double y, double z,
                    double t, emf_t *o)
{
    o->Bx = o->By = o->Bz = o->Ex = o->Ey = o->Ez = 0.0;
    switch (f->kind) {
    case TPREF_FIELD_TEST:
        o->Bz = 1.0;
        break;
    case TPREF_FIELD_ABC: {
        /* getb_ff, src/force_free.c:138-150 */
        double A = f->abc[0], B = f->abc[1], C = f->abc[2];
        double lam = f->abc[3], B0 = f->abc[4];
        double bx = A * o_sin(lam * z) + C * o_cos(lam * y);
        double by = B * o_sin(lam * x) + A * o_cos(lam * z);
        double bz = C * o_sin(lam * y) + B * o_cos(lam * x);
        bx *= B0; by *= B0; bz *= B0;
        o->Bx = bx; o->By = by; o->Bz = bz;
        if (f->vgrid) {
            /* gete_ff, src/force_free.c:78-86 */
            long c[8];
            double w[8], v[3];
            trilinear_ref(f->vnx, f->vny, f->vnz, f->vlo, f->vhi, x, y, z, c, w);
            gather_double(NULL, f->vgrid, c, w, v);
            o->Ex = by * v[2] - bz * v[1];
            o->Ey = bz * v[0] - bx * v[2];
            o->Ez = bx * v[1] - by * v[0];
        } else {
            o->Ex = o->Ey = o->Ez = 0.0;
        }
        break;
    }
    case TPREF_FIELD_GRID: {
        /* getemf_pic_test_particle, src/emfields.c:153-166 */
        long c[8];
        double w[8];
        if (f->egrid_b) {
            /* two time slices (multi_tframe = 1): get_float_efield_at_point
             * / get_float_bfield_at_point (src/electric_field.c:407-437,
             * src/magnetic_field.c:593-622): each slice summed in double,
             * blended dt * next + (1 - dt) * previous with dt the fraction of
             * t / grid_dt (src/interpolation.c:97-100), returned as float
             * (the textbook mode keeps the double) */
            double a[6], b[6], v[6];
            if (f->gather_mode == TP_GATHER_REF_COMPAT)
                trilinear_ref(f->nx, f->ny, f->nz, f->lo, f->hi, x, y, z, c, w);
            else
                trilinear_textbook(f->nx, f->ny, f->nz, f->lo, f->hi, x, y, z, c, w);
            gather_double(f->egrid, NULL, c, w, a);
            gather_double(f->bgrid, NULL, c, w, a + 3);
            gather_double(f->egrid_b, NULL, c, w, b);
            gather_double(f->bgrid_b, NULL, c, w, b + 3);
            double r = t / f->grid_dt;
            double fl = floor(r);
            int it1 = (fl > -2147483649.0 && fl < 2147483648.0) ? (int)fl : (int)0x80000000;
            double dtf = r - it1;
            for (int k = 0; k < 6; k++) v[k] = dtf * a[k] + (1.0 - dtf) * b[k];
            if (f->gather_mode == TP_GATHER_REF_COMPAT)
                for (int k = 0; k < 6; k++) v[k] = (double)(float)v[k];
            o->Ex = v[0]; o->Ey = v[1]; o->Ez = v[2];
            o->Bx = v[3]; o->By = v[4]; o->Bz = v[5];
        } else if (f->gather_mode == TP_GATHER_REF_COMPAT) {
            float e[3], b[3];
            trilinear_ref(f->nx, f->ny, f->nz, f->lo, f->hi, x, y, z, c, w);
            gather_float_ref(f->egrid, c, w, e);
            gather_float_ref(f->bgrid, c, w, b);
            o->Ex = e[0]; o->Ey = e[1]; o->Ez = e[2];
            o->Bx = b[0]; o->By = b[1]; o->Bz = b[2];
        } else {
            double e[3], b[3];
            trilinear_textbook(f->nx, f->ny, f->nz, f->lo, f->hi, x, y, z, c, w);
            gather_double(f->egrid, NULL, c, w, e);
            gather_double(f->bgrid, NULL, c, w, b);
            o->Ex = e[0]; o->Ey = e[1]; o->Ez = e[2];
            o->Bx = b[0]; o->By = b[1]; o->Bz = b[2];
        }
        break;
    }
    case TPREF_FIELD_MHD: {
        /* getemf_mhd_test_particle, src/emfields.c:107-120: the float
         * velocity and magnetic gathers (get_float_velocity_at_point,
         * src/velocity_field.c:593-622; get_float_bfield_at_point,
         * src/magnetic_field.c:593-622), E = B x v in double */
        long c[8];
        double w[8];
        float v[3], b[3];
        trilinear_ref(f->nx, f->ny, f->nz, f->lo, f->hi, x, y, z, c, w);
        gather_float_ref(f->egrid, c, w, v);
        gather_float_ref(f->bgrid, c, w, b);
        o->Bx = b[0]; o->By = b[1]; o->Bz = b[2];
        o->Ex = o->By * v[2] - o->Bz * v[1];
        o->Ey = o->Bz * v[0] - o->Bx * v[2];
        o->Ez = o->Bx * v[1] - o->By * v[0];
        break;
    }
    case TPREF_FIELD_GRIDD: {
        /* get_double_efield_at_point / get_double_bfield_at_point
         * (src/electric_field.c:370-406, src/magnetic_field.c:555-591) */
        long c[8];
        double w[8], e[3], b[3];
        trilinear_ref(f->nx, f->ny, f->nz, f->lo, f->hi, x, y, z, c, w);
        gather_double(NULL, f->egrid_d, c, w, e);
        gather_double(NULL, f->bgrid_d, c, w, b);
        o->Ex = e[0]; o->Ey = e[1]; o->Ez = e[2];
        o->Bx = b[0]; o->By = b[1]; o->Bz = b[2];
        break;
    }
    case TPREF_FIELD_MODES: {
        /* B = b0z z + sum_n a_n cos(k_n.x + phi_n); the phase is built
         * as fma(kz, z, fma(ky, y, fma(kx, x, phi))) and the sum uses fma */
        double bx = 0.0, by = 0.0, bz = 0.0;
        for (int n = 0; n < f->nmodes; n++) {
            const double *m = f->modes + 7 * n;
            double ph = fma(m[2], z, fma(m[1], y, fma(m[0], x, m[6])));
            double cs = g_use_libm ? cos(ph) : tp_cos_pi(ph);
            bx = fma(m[3], cs, bx);
            by = fma(m[4], cs, by);
            bz = fma(m[5], cs, bz);
        }
        o->Bx = bx; o->By = by; o->Bz = f->b0z + bz;
        break;
    }
    case TPREF_FIELD_WLCS:
        emf_wlcs(f, x, y, z, t, o);
        break;
    default:
        break;
    }
}

void tpref_eval_field(const tpref_field *f, int n, const double *pts,
                      double *out)
{
    for (int i = 0; i < n; i++) {
        emf_t e;
        get_emf(f, pts[4 * i], pts[4 * i + 1], pts[4 * i + 2], pts[4 * i + 3], &e);
        double *o = out + 6 * i;
        o[0] = e.Bx; o[1] = e.By; o[2] = e.Bz;
        o[3] = e.Ex; o[4] = e.Ey; o[5] = e.Ez;
    }
}

/* ---------------------------------------------------------------------- */
/* pushers                                                                 */
/* ---------------------------------------------------------------------- */
static inline double modulus(double x, double y, double z)
{
    return sqrt(x * x + y * y + z * z);
}

/* tracking_wirz, src/tracking.c:749-892, same operation order */
void tpref_wirz_step(const tpref_field *f, tp_particle *p, double dt, double qm)
{
    double vminus[3], vplus[3], delta[3], delta_m[3], hpara[3], hperp[3], hr[3];
    double r0[3] = {p->x, p->y, p->z};
    double t0 = p->t;
    emf_t emf;
    delta[0] = p->ux; delta[1] = p->uy; delta[2] = p->uz;
    double gama, igama, Btot, iBtot, vdothp, vperp, gyroR, omegac, dtheta;
    double sd, cd, tmp1, tmp2, tmp3;
    int i;

    get_emf(f, r0[0], r0[1], r0[2], t0, &emf);
    Btot = sqrt(emf.Bx * emf.Bx + emf.By * emf.By + emf.Bz * emf.Bz);
    iBtot = 1.0 / Btot;
    double dth = dt * 0.5;
    delta_m[0] = delta[0] + qm * dth * emf.Ex;
    delta_m[1] = delta[1] + qm * dth * emf.Ey;
    delta_m[2] = delta[2] + qm * dth * emf.Ez;
    gama = sqrt(1.0 + delta_m[0] * delta_m[0] + delta_m[1] * delta_m[1] +
                delta_m[2] * delta_m[2]);
    igama = 1.0 / gama;
    for (i = 0; i < 3; i++) vminus[i] = delta_m[i] * igama;
    hpara[0] = emf.Bx * iBtot;
    hpara[1] = emf.By * iBtot;
    hpara[2] = emf.Bz * iBtot;
    vdothp = 0.0;
    for (i = 0; i < 3; i++) vdothp += vminus[i] * hpara[i];
    vperp = modulus(vminus[0] - vdothp * hpara[0], vminus[1] - vdothp * hpara[1],
                    vminus[2] - vdothp * hpara[2]);
    for (i = 0; i < 3; i++) hperp[i] = (vminus[i] - vdothp * hpara[i]) / vperp;
    hr[0] = hpara[1] * hperp[2] - hpara[2] * hperp[1];
    hr[1] = hpara[2] * hperp[0] - hpara[0] * hperp[2];
    hr[2] = hpara[0] * hperp[1] - hpara[1] * hperp[0];
    gyroR = vperp * gama * 1.0 * iBtot / qm;
    omegac = qm * Btot * igama;
    dtheta = omegac * dt;
    sd = o_sin(dtheta * 0.5);
    cd = o_cos(dtheta * 0.5);
    tmp1 = dth * vdothp * 1.0;
    tmp2 = gyroR * sd;
    tmp3 = -gyroR * (1.0 - cd);
    double pre[3];
    for (i = 0; i < 3; i++)
        pre[i] = r0[i] + hpara[i] * tmp1 + hperp[i] * tmp2 + hr[i] * tmp3;

    get_emf(f, pre[0], pre[1], pre[2], t0 + dth, &emf);
    Btot = sqrt(emf.Bx * emf.Bx + emf.By * emf.By + emf.Bz * emf.Bz);
    iBtot = 1.0 / Btot;
    hpara[0] = emf.Bx * iBtot;
    hpara[1] = emf.By * iBtot;
    hpara[2] = emf.Bz * iBtot;
    vdothp = 0.0;
    for (i = 0; i < 3; i++) vdothp += vminus[i] * hpara[i];
    vperp = modulus(vminus[0] - vdothp * hpara[0], vminus[1] - vdothp * hpara[1],
                    vminus[2] - vdothp * hpara[2]);
    for (i = 0; i < 3; i++) hperp[i] = (vminus[i] - vdothp * hpara[i]) / vperp;
    hr[0] = hpara[1] * hperp[2] - hpara[2] * hperp[1];
    hr[1] = hpara[2] * hperp[0] - hpara[0] * hperp[2];
    hr[2] = hpara[0] * hperp[1] - hpara[1] * hperp[0];
    omegac = qm * Btot * igama;
    dtheta = omegac * dt;
    sd = o_sin(dtheta);
    cd = o_cos(dtheta);
    for (i = 0; i < 3; i++)
        vplus[i] = hpara[i] * vdothp + vperp * (cd * hperp[i] - sd * hr[i]);
    /* get_delta_vec, src/tracking.c:1066-1078 */
    double bm = 0.0;
    for (i = 0; i < 3; i++) bm += vplus[i] * vplus[i];
    double g2 = 1.0 / sqrt(1.0 - bm);
    for (i = 0; i < 3; i++) delta[i] = g2 * vplus[i];
    p->ux = delta[0] + qm * dth * emf.Ex;
    p->uy = delta[1] + qm * dth * emf.Ey;
    p->uz = delta[2] + qm * dth * emf.Ez;
    vdothp = 0.0;
    for (i = 0; i < 3; i++) vdothp += vplus[i] * hpara[i];
    vperp = modulus(vplus[0] - vdothp * hpara[0], vplus[1] - vdothp * hpara[1],
                    vplus[2] - vdothp * hpara[2]);
    gyroR = vperp * gama * 1.0 * iBtot / qm;
    tmp1 = dt * vdothp * 1.0;
    tmp2 = gyroR * sd;
    tmp3 = -gyroR * (1.0 - cd);
    double d[3];
    for (i = 0; i < 3; i++) d[i] = hpara[i] * tmp1 + hperp[i] * tmp2 + hr[i] * tmp3;
    p->x += d[0];
    p->y += d[1];
    p->z += d[2];
    p->t += dt;
}

/* Relativistic Boris push (not in the reference; it occupies the call slot
 * of src/tracking.c:429).  u = gamma v; field at (x, t):
 *   h = qm dt/2;  u- = u + h E;  g- = sqrt(1 + |u-|^2);  tv = (h/g-) B;
 *   sv = 2 tv / (1 + |tv|^2);  u' = u- + u- x tv;  u+ = u- + u' x sv;
 *   u = u+ + h E;  x += u (dt / sqrt(1 + |u|^2));  t += dt */
void tpref_boris_step(const tpref_field *f, tp_particle *p, double dt, double qm)
{
    emf_t e;
    get_emf(f, p->x, p->y, p->z, p->t, &e);
    double h = qm * (0.5 * dt);
    double umx = p->ux + h * e.Ex;
    double umy = p->uy + h * e.Ey;
    double umz = p->uz + h * e.Ez;
    double gm = sqrt(1.0 + umx * umx + umy * umy + umz * umz);
    double fq = h / gm;
    double tx = fq * e.Bx, ty = fq * e.By, tz = fq * e.Bz;
    double t2 = tx * tx + ty * ty + tz * tz;
    double sf = 2.0 / (1.0 + t2);
    double sx = sf * tx, sy = sf * ty, sz = sf * tz;
    double upx = umx + (umy * tz - umz * ty);
    double upy = umy + (umz * tx - umx * tz);
    double upz = umz + (umx * ty - umy * tx);
    double uqx = umx + (upy * sz - upz * sy);
    double uqy = umy + (upz * sx - upx * sz);
    double uqz = umz + (upx * sy - upy * sx);
    double ux = uqx + h * e.Ex;
    double uy = uqy + h * e.Ey;
    double uz = uqz + h * e.Ez;
    double g = sqrt(1.0 + ux * ux + uy * uy + uz * uz);
    double ig = dt / g;
    p->ux = ux; p->uy = uy; p->uz = uz;
    p->x += ux * ig;
    p->y += uy * ig;
    p->z += uz * ig;
    p->t += dt;
}

/* particle_bc, src/tracking.c:1172-1221 */
void tpref_particle_bc(const tp_track_params *q, tp_particle *p, int *iescape)
{
    double xd = q->xmax - q->xmin, yd = q->ymax - q->ymin, zd = q->zmax - q->zmin;
    if (q->bc_flag == 0) {
        if (p->x > q->xmax) { p->x -= xd; p->xshift += xd; }
        if (p->y > q->ymax) { p->y -= yd; p->yshift += yd; }
        if (p->z > q->zmax) { p->z -= zd; p->zshift += zd; }
        if (p->x < q->xmin) { p->x += xd; p->xshift -= xd; }
        if (p->y < q->ymin) { p->y += yd; p->yshift -= yd; }
        if (p->z < q->zmin) { p->z += zd; p->zshift -= zd; }
    } else if (q->bc_flag == 1) {
        if (p->x > q->xmax || p->y > q->ymax || p->z > q->zmax ||
            p->x < q->xmin || p->y < q->ymin || p->z < q->zmin)
            *iescape = 1;
    } else if (q->bc_flag == 2) {
        *iescape = 0;
    }
}

/* ptl_energy_fixed, src/diagnostics.c:187-205.  The int conversion of
 * floor() follows x86-64: out-of-range and NaN give INT_MIN, hence bin 0. */
int tpref_energy_bin(const tp_particle *p, int nbins, double logemin, double logde)
{
    double gama = sqrt(1.0 + p->ux * p->ux + p->uy * p->uy + p->uz * p->uz);
    double ene = gama - 1.0;
    double v = floor((log10(ene) - logemin) / logde);
    if (!(v > -2147483649.0 && v < 2147483648.0)) return 0;
    int ibin = (int)v;
    if (ibin < 0) ibin = 0;
    else if (ibin > nbins - 1) ibin = nbins - 1;
    return ibin;
}

/* calc_diff_coeff, src/diagnostics.c:659-705 */
void tpref_diff_coeff(const tpref_field *f, const tp_particle *p0,
                      const tp_particle *p, double dt, double o[8])
{
    double t = p->t * 2 + dt;
    double a;
    a = p0->x - (p->x + p->xshift); double dxx = a * a / t;
    a = p0->y - (p->y + p->yshift); double dyy = a * a / t;
    a = p0->z - (p->z + p->zshift); double dzz = a * a / t;
    double drr = dxx + dyy + dzz;
    double ax = p0->ux - p->ux, ay = p0->uy - p->uy, az = p0->uz - p->uz;
    double dpp = (ax * ax + ay * ay + az * az) / t;
    double g = sqrt(1.0 + p->ux * p->ux + p->uy * p->uy + p->uz * p->uz);
    double g0 = sqrt(1.0 + p0->ux * p0->ux + p0->uy * p0->uy + p0->uz * p0->uz);
    a = g - g0; double dee = a * a / t;
    emf_t e, e0;
    get_emf(f, p->x, p->y, p->z, p->t, &e);
    get_emf(f, p0->x, p0->y, p0->z, p0->t, &e0);
    double ab = sqrt(e.Bx * e.Bx + e.By * e.By + e.Bz * e.Bz);
    double au = sqrt(p->ux * p->ux + p->uy * p->uy + p->uz * p->uz);
    double mu = (e.Bx * p->ux + e.By * p->uy + e.Bz * p->uz) / (ab * au);
    ab = sqrt(e0.Bx * e0.Bx + e0.By * e0.By + e0.Bz * e0.Bz);
    au = sqrt(p0->ux * p0->ux + p0->uy * p0->uy + p0->uz * p0->uz);
    double mu0 = (e0.Bx * p0->ux + e0.By * p0->uy + e0.Bz * p0->uz) / (ab * au);
    a = mu - mu0; double duu = a * a / t;
    double th = acos(mu) * 180 / M_PI;
    double th0 = acos(mu0) * 180 / M_PI;
    a = th - th0; double daa = a * a / t;
    o[0] = drr; o[1] = dxx; o[2] = dyy; o[3] = dzz;
    o[4] = dpp; o[5] = duu; o[6] = daa; o[7] = dee;
}

/* ---------------------------------------------------------------------- */
/* fixed-step loop, src/tracking.c:90-190 and :297-462                     */
/* ---------------------------------------------------------------------- */
/* Trajectory diagnostics of the next tpref_track_fixed call (the
 * reference's traj_diagnose = 1, src/tracking.c:440-447): particle i is
 * copied into ptl_time[ntraj_accum[i-1] + k], k = 1, 2, ..., at every step
 * j with j % nsteps_output == 0 at which it is tracked; points[i] receives
 * the final k (starting at 1: slot 0 is the caller's initial point).  The
 * reference does not bound k; here a point past the particle's slots
 * (ntraj_accum[i] - ntraj_accum[i-1]) is counted but not written.
 * nsteps_output = 0 switches the diagnostics off. */
static int g_traj_nout = 0;
static const int *g_traj_accum = NULL;
static tp_particle *g_traj = NULL;
static int *g_traj_points = NULL;

void tpref_set_trajectory(int nsteps_output, const int *ntraj_accum, tp_particle *ptl_time,
                          int *points)
{
    g_traj_nout = nsteps_output;
    g_traj_accum = ntraj_accum;
    g_traj = ptl_time;
    g_traj_points = points;
}

int tpref_track_fixed(const tpref_field *f, const tp_track_params *q,
                      int n, tp_particle *ptl, const tp_particle *ptl_init,
                      int *nsteps, double *espect, double *espect_escape,
                      double *dcoeff, int *nptl_remain, int nthreads,
                      double *seconds)
{
    if (n < 0 || q->einterval_t <= 0 || q->dcoeff_interval <= 0 ||
        q->estep < q->sstep || q->nbins < 1 || q->nt_out < 1)
        return TP_E_ARG;
    const int nb = q->nbins, nt = q->nt_out, di = q->dcoeff_interval;
    const int nd = q->estep / di + 1;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    /* per-thread accumulators, merged in thread order at the end */
    double *dsum = calloc((size_t)nthreads * 8 * nd, sizeof(double));
    int *rem = calloc((size_t)nthreads * nd, sizeof(int));
    double *esp = calloc((size_t)nthreads * nb * nt, sizeof(double));
    double *ese = calloc((size_t)nthreads * nb * nt, sizeof(double));
    if (!dsum || !rem || !esp || !ese) {
        free(dsum); free(rem); free(esp); free(ese);
        return TP_E_NOMEM;
    }
    double t1 = omp_get_wtime();
#pragma omp parallel num_threads(nthreads)
    {
        int tid = omp_get_thread_num();
        double *ds = dsum + (size_t)tid * 8 * nd;
        int *rm = rem + (size_t)tid * nd;
        double *sp = esp + (size_t)tid * nb * nt;
        double *se = ese + (size_t)tid * nb * nt;
        double dcarry[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma omp for schedule(static)
        for (int i = 0; i < n; i++) {
            tp_particle *p = &ptl[i];
            int esc_pre = 0, esc_after;
            if (nsteps) nsteps[i] = 1;
            tpref_particle_bc(q, p, &esc_pre);
            esc_after = esc_pre;
            int ntp = (int)floor(q->sstep / (double)q->einterval_t);
            int tstep = 1;
            /* the reference's stale per-particle terms start undefined;
             * here they start at zero */
            double dv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (g_carry_stale) memcpy(dv, dcarry, sizeof(dv));
            if (esc_pre && 1 < nd) rm[1]--;
            const int toff = (g_traj_nout > 0 && i > 0) ? g_traj_accum[i - 1] : 0;
            const int tcap = g_traj_nout > 0 ? g_traj_accum[i] - toff : 0;
            int tpts = 1;
            for (int j = q->sstep; j <= q->estep; j++) {
                dv[0] = 0.0; dv[4] = 0.0; dv[5] = 0.0;  /* drr dpp duu */
                if (q->flags & TP_FLAG_FIX_STALE_DCOEFF)
                    for (int k = 0; k < 8; k++) dv[k] = 0.0;
                if (esc_pre == 0 || q->bc_flag == 0) {
                    if (q->pusher == TP_PUSH_WIRZ) tpref_wirz_step(f, p, q->dt, q->qm);
                    else tpref_boris_step(f, p, q->dt, q->qm);
                    tpref_particle_bc(q, p, &esc_after);
                    if (nsteps) nsteps[i] += 1;
                    if (esc_after == 0 && j % di == 0)
                        tpref_diff_coeff(f, &ptl_init[i], p, q->dt, dv);
                    if (g_traj_nout > 0 && j % g_traj_nout == 0) {
                        if (tpts < tcap) memcpy(&g_traj[toff + tpts], p, sizeof(tp_particle));
                        tpts++;
                    }
                }
                if (esc_pre == 0 && esc_after == 1 && q->bc_flag == 1 && ntp < nt)
                    se[ntp * nb + tpref_energy_bin(p, nb, q->logemin, q->logde)] += 1.0;
                if (j % q->einterval_t == 0) {
                    if (esc_after == 0 && ntp < nt)
                        sp[ntp * nb + tpref_energy_bin(p, nb, q->logemin, q->logde)] += 1.0;
                    ntp++;
                }
                if (j % di == 0) {
                    if (tstep < nd)
                        for (int k = 0; k < 8; k++) ds[k * nd + tstep] += dv[k];
                    tstep++;
                }
                if (esc_after && j % di == 0 && tstep < nd) rm[tstep]--;
                esc_pre = esc_after;
            }
            memcpy(dcarry, dv, sizeof(dv));
            if (g_traj_nout > 0 && g_traj_points) g_traj_points[i] = tpts;
        }
    }
    double t2 = omp_get_wtime();
    if (seconds) *seconds = t2 - t1;
    if (dcoeff) {
        for (int k = 0; k < 8 * nd; k++) {
            double s = 0.0;
            for (int th = 0; th < nthreads; th++) s += dsum[(size_t)th * 8 * nd + k];
            dcoeff[k] = s;
        }
    }
    if (nptl_remain) {
        for (int k = 0; k < nd; k++) {
            int s = n;
            for (int th = 0; th < nthreads; th++) s += rem[(size_t)th * nd + k];
            nptl_remain[k] = s;
        }
    }
    for (int k = 0; k < nb * nt; k++) {
        double a = 0.0, b = 0.0;
        for (int th = 0; th < nthreads; th++) {
            a += esp[(size_t)th * nb * nt + k];
            b += ese[(size_t)th * nb * nt + k];
        }
        if (espect) espect[k] = a;
        if (espect_escape) espect_escape[k] = b;
    }
    free(dsum); free(rem); free(esp); free(ese);
    return TP_OK;
}


/* ---------------------------------------------------------------------- */
/* adaptive Bulirsch-Stoer (algorithm of src/unused/StepperBS.c:103-644,   */
/* NR3 ch. 17, dense output optional), per-particle controller state,      */
/* ODE in (x, u) with du/dt = qm (E + u/gamma x B).                        */
/* ---------------------------------------------------------------------- */
#define BS_N 6
#define BS_STEPFAC1 0.65
#define BS_STEPFAC2 0.94
#define BS_STEPFAC3 0.02
#define BS_STEPFAC4 4.0
#define BS_KFAC1 0.8
#define BS_KFAC2 0.9

typedef struct {
    int kmax, ktarg0, dense;
    int nseq[10], cost[10], ipoint[11];
    double coeff[10][10];
    double errfac[19];
} bs_consts;

/* StepperBase, src/unused/StepperBS.c:103-149: the step-size sequence
 * (dense output: nseq = 4k+2, else 2(k+1)), costs, k_targ from rtol, the
 * extrapolation coefficients, the dense interpolation error factors and
 * ipoint, where each sequence's derivatives start in fsave */
static void bs_setup(int kmax, double rtol, int dense, bs_consts *c)
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

static void bs_rhs(const tpref_field *f, double qm, double t, const double *y, double *dy)
{
    emf_t e;
    get_emf(f, y[0], y[1], y[2], t, &e);
    double g = sqrt(1.0 + y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
    double ig = 1.0 / g;
    double vx = y[3] * ig, vy = y[4] * ig, vz = y[5] * ig;
    dy[0] = vx;
    dy[1] = vy;
    dy[2] = vz;
    dy[3] = qm * (e.Ex + (vy * e.Bz - vz * e.By));
    dy[4] = qm * (e.Ey + (vz * e.Bx - vx * e.Bz));
    dy[5] = qm * (e.Ez + (vx * e.By - vy * e.Bx));
}

/* dense-output storage of one particle (StepperBS's ysave, fsave, dens) */
#define BS_FSAVE 162            /* ipoint[9] = 2 * 9^2 for kmax = 8 */
typedef struct {
    double ysave[9][BS_N];
    double fsave[BS_FSAVE][BS_N];
    double dens[(2 * 9 + 5) * BS_N];
    int mu;
    double xold, h;             /* the step the polynomial covers */
} bs_dense;

typedef struct {
    double t, hnext;
    double dydx[BS_N];
    int ktarg, first, last, reject, prev_reject;
    long long rhs;
    int acc, rej, attempts;
} bs_state;

/* updatey, src/unused/StepperBS.c:364-414: the modified midpoint rule over
 * nstep substeps.  With dense output the state at substep nstep/2 goes to
 * ysave[k] and every substep's derivative (substeps 1 .. nstep) to fsave,
 * consecutive rows from *ipt on (the reference passes ipt by value, so its
 * rows would all start at 0; NR3's dy passes it by reference, as here). */
static void bs_midpoint(const tpref_field *f, double qm, double x, const double *y,
                        const double *dydx, double htot, int k, int nstep, double *yend,
                        long long *rhs, bs_dense *d, int *ipt)
{
    double ym[BS_N], yn[BS_N];
    double h = htot / nstep;
    for (int i = 0; i < BS_N; i++) {
        ym[i] = y[i];
        yn[i] = y[i] + h * dydx[i];
    }
    double xnew = x + h;
    bs_rhs(f, qm, xnew, yn, yend);
    (*rhs)++;
    double h2 = 2.0 * h;
    for (int nn = 1; nn < nstep; nn++) {
        if (d && nn == nstep / 2)
            for (int i = 0; i < BS_N; i++) d->ysave[k][i] = yn[i];
        if (d && abs(nn - nstep / 2) <= 2 * k + 1) {
            (*ipt)++;
            for (int i = 0; i < BS_N; i++) d->fsave[*ipt][i] = yend[i];
        }
        for (int i = 0; i < BS_N; i++) {
            double swap = ym[i] + h2 * yend[i];
            ym[i] = yn[i];
            yn[i] = swap;
        }
        xnew += h;
        bs_rhs(f, qm, xnew, yn, yend);
        (*rhs)++;
    }
    if (d && nstep / 2 <= 2 * k + 1) {
        (*ipt)++;
        for (int i = 0; i < BS_N; i++) d->fsave[*ipt][i] = yend[i];
    }
    for (int i = 0; i < BS_N; i++) yend[i] = 0.5 * (ym[i] + yn[i] + h * yend[i]);
}

/* dense_interp, src/unused/StepperBS.c:478-524: the coefficients of the
 * interpolating polynomial from the values prepare_dense stored in y */
static void bs_dense_interp(double *y, int imit)
{
    double a[31];
    for (int i = 0; i < BS_N; i++) {
        const int n = BS_N;
        double y0 = y[i], y1 = y[2 * n + i], yp0 = y[n + i], yp1 = y[3 * n + i];
        double ydiff = y1 - y0;
        double aspl = -yp1 + ydiff;
        double bspl = yp0 - ydiff;
        y[n + i] = ydiff;
        y[2 * n + i] = aspl;
        y[3 * n + i] = bspl;
        if (imit < 0) continue;
        double ph0 = (y0 + y1) * 0.5 + 0.125 * (aspl + bspl);
        double ph1 = ydiff + (aspl - bspl) * 0.25;
        double ph2 = -(yp0 - yp1);
        double ph3 = 6.0 * (bspl - aspl);
        if (imit >= 1) {
            a[1] = 16.0 * (y[5 * n + i] - ph1);
            if (imit >= 3) {
                a[3] = 16.0 * (y[7 * n + i] - ph3 + 3 * a[1]);
                for (int im = 5; im <= imit; im += 2) {
                    double fac1 = im * (im - 1) / 2.0;
                    double fac2 = fac1 * (im - 2) * (im - 3) * 2.0;
                    a[im] = 16.0 * (y[(im + 4) * n + i] + fac1 * a[im - 2] - fac2 * a[im - 4]);
                }
            }
        }
        a[0] = (y[4 * n + i] - ph0) * 16.0;
        if (imit >= 2) {
            a[2] = (y[n * 6 + i] - ph2 + a[0]) * 16.0;
            for (int im = 4; im <= imit; im += 2) {
                double fac1 = im * (im - 1) / 2.0;
                double fac2 = (double)(im * (im - 1) * (im - 2) * (im - 3));
                a[im] = 16.0 * (y[(im + 4) * n + i] + fac1 * a[im - 2] - fac2 * a[im - 4]);
            }
        }
        for (int im = 0; im <= imit; im++) y[n * (im + 4) + i] = a[im];
    }
}

/* prepare_dense, src/unused/StepperBS.c:526-615: the dense-output values at
 * the step of size h (ysav at its start, y and dydxnew at its end, dydx the
 * derivative at the start), extrapolated from the midpoint values and
 * derivative differences of the first k+1 sequences; returns the
 * interpolation error estimate */
static double bs_prepare_dense(const bs_consts *c, double h, const double *dydxnew,
                               const double *ysav, const double *scale, int k,
                               const double *y, const double *dydx, bs_dense *d)
{
    const int n = BS_N;
    const int mu = 2 * k - 1;
    double *dens = d->dens;
    for (int i = 0; i < n; i++) {
        dens[i] = ysav[i];
        dens[n + i] = h * dydx[i];
        dens[2 * n + i] = y[i];
        dens[3 * n + i] = dydxnew[i] * h;
    }
    for (int j = 1; j <= k; j++) {
        double dblenj = c->nseq[j];
        for (int l = j; l >= 1; l--) {
            double r = dblenj / c->nseq[l - 1];
            double factor = r * r - 1.0;
            for (int i = 0; i < n; i++)
                d->ysave[l - 1][i] = d->ysave[l][i] + (d->ysave[l][i] - d->ysave[l - 1][i]) / factor;
        }
    }
    for (int i = 0; i < n; i++) dens[4 * n + i] = d->ysave[0][i];
    for (int kmi = 1; kmi <= mu; kmi++) {
        int kbeg = (kmi - 1) / 2;
        for (int kk = kbeg; kk <= k; kk++) {
            double facnj = tp_powi(c->nseq[kk] / 2.0, kmi - 1);
            int ipt = c->ipoint[kk + 1] - 2 * kk + kmi - 3;
            for (int i = 0; i < n; i++) d->ysave[kk][i] = d->fsave[ipt][i] * facnj;
        }
        for (int j = kbeg + 1; j <= k; j++) {
            double dblenj = c->nseq[j];
            for (int l = j; l >= kbeg + 1; l--) {
                double r = dblenj / c->nseq[l - 1];
                double factor = r * r - 1.0;
                for (int i = 0; i < n; i++)
                    d->ysave[l - 1][i] =
                        d->ysave[l][i] + (d->ysave[l][i] - d->ysave[l - 1][i]) / factor;
            }
        }
        for (int i = 0; i < n; i++) dens[(kmi + 4) * n + i] = d->ysave[kbeg][i] * h;
        if (kmi == mu) continue;
        for (int kk = kmi / 2; kk <= k; kk++) {
            int lbeg = c->ipoint[kk + 1] - 1;
            int lend = c->ipoint[kk] + kmi;
            if (kmi == 1) lend += 2;
            for (int l = lbeg; l >= lend; l -= 2)
                for (int i = 0; i < n; i++) d->fsave[l][i] = d->fsave[l][i] - d->fsave[l - 2][i];
            if (kmi == 1) {
                int l = lend - 2;
                for (int i = 0; i < n; i++) d->fsave[l][i] = d->fsave[l][i] - dydx[i];
            }
        }
        for (int kk = kmi / 2; kk <= k; kk++) {
            int lbeg = c->ipoint[kk + 1] - 2;
            int lend = c->ipoint[kk] + kmi + 1;
            for (int l = lbeg; l >= lend; l -= 2)
                for (int i = 0; i < n; i++) d->fsave[l][i] = d->fsave[l][i] - d->fsave[l - 2][i];
        }
    }
    bs_dense_interp(dens, mu);
    d->mu = mu;
    double error = 0.0;
    if (mu >= 1) {
        for (int i = 0; i < n; i++) {
            double q = dens[(mu + 4) * n + i] / scale[i];
            error += q * q;
        }
        error = sqrt(error / n) * c->errfac[mu - 1];
    }
    return error;
}

/* dense_out, src/unused/StepperBS.c:617-644: component i at time x in the
 * step [xold, xold + h], with the polynomial of the last prepare_dense */
static double bs_dense_out(const bs_dense *d, int i, double x)
{
    const int n = BS_N, mu = d->mu;
    const double *dens = d->dens;
    double theta = (x - d->xold) / d->h;
    double theta1 = 1.0 - theta;
    double yinterp = dens[i] + theta * (dens[n + i] + theta1 * (dens[2 * n + i] * theta +
                                                                dens[3 * n + i] * theta1));
    if (mu < 0) return yinterp;
    double theta05 = theta - 0.5;
    double tt = theta * theta1;
    double t4 = tt * tt;
    double cc = dens[n * (mu + 4) + i];
    for (int j = mu; j > 0; j--) cc = dens[n * (j + 3) + i] + cc * theta05 / j;
    return yinterp + t4 * cc;
}

/* step, src/unused/StepperBS.c:159-354.  With d (dense output) the accepted
 * step's interpolation error is estimated by prepare_dense; above 10 the
 * step is redone with the step size it suggests, and the next step size is
 * capped by it.  Returns 0, or 1 when the attempt cap was reached. */
static int bs_step(const tpref_field *f, const tp_bs_params *prm, const bs_consts *c,
                   bs_state *st, double *y, double htry, int cap, bs_dense *d)
{
    double table[9][BS_N], hopt[10], work[10];
    double ysav[BS_N], yseq[BS_N], scale[BS_N];
    double h = htry, hnew, err = 0.0, hopt_int = 0.0;
    int k = 0;
    work[0] = 0.0;
    for (int i = 0; i < BS_N; i++) ysav[i] = y[i];
    if (h != st->hnext && !st->first) st->last = 1;
    if (st->reject) {
        st->prev_reject = 1;
        st->last = 0;
    }
    st->reject = 0;
    int firstk = 1;
    hnew = fabs(h);
    double dnew[BS_N];
    for (;;) {
    while (firstk || st->reject) {
        h = hnew;
        firstk = 0;
        st->reject = 0;
        if (++st->attempts > cap) {
            for (int i = 0; i < BS_N; i++) y[i] = ysav[i];
            return 1;
        }
        int ipt = -1;
        for (k = 0; k <= st->ktarg + 1; k++) {
            bs_midpoint(f, prm->qm, st->t, ysav, st->dydx, h, k, c->nseq[k], yseq, &st->rhs,
                        d, &ipt);
            if (k == 0) {
                for (int i = 0; i < BS_N; i++) y[i] = yseq[i];
                continue;
            }
            for (int i = 0; i < BS_N; i++) table[k - 1][i] = yseq[i];
            /* polyextr, :419-435 */
            for (int j = k - 1; j > 0; j--)
                for (int i = 0; i < BS_N; i++)
                    table[j - 1][i] = table[j][i] + c->coeff[k][j] * (table[j][i] - table[j - 1][i]);
            for (int i = 0; i < BS_N; i++)
                y[i] = table[0][i] + c->coeff[k][0] * (table[0][i] - y[i]);
            err = 0.0;
            for (int i = 0; i < BS_N; i++) {
                double a = fabs(ysav[i]), b = fabs(y[i]);
                scale[i] = prm->atol + prm->rtol * (a > b ? a : b);
                double q = (y[i] - table[0][i]) / scale[i];
                err += q * q;
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
            hopt[k] = fabs(h * fac);
            work[k] = c->cost[k] / hopt[k];
            if ((st->first || st->last) && err <= 1.0) break;
            if (k == st->ktarg - 1 && !st->prev_reject && !st->first && !st->last) {
                if (err <= 1.0) break;
                double r = (double)(c->nseq[st->ktarg] * c->nseq[st->ktarg + 1] /
                                    (c->nseq[0] * c->nseq[0]));
                if (err > r * r) {
                    st->reject = 1;
                    st->ktarg = k;
                    if (st->ktarg > 1 && work[k - 1] < BS_KFAC1 * work[k]) st->ktarg--;
                    hnew = hopt[st->ktarg];
                    break;
                }
            }
            if (k == st->ktarg) {
                if (err <= 1.0) break;
                double r = (double)c->nseq[k + 1] / c->nseq[0];
                if (err > r * r) {
                    st->reject = 1;
                    if (st->ktarg > 1 && work[k - 1] < BS_KFAC1 * work[k]) st->ktarg--;
                    hnew = hopt[st->ktarg];
                    break;
                }
            }
            if (k == st->ktarg + 1) {
                if (err > 1.0) {
                    st->reject = 1;
                    if (st->ktarg > 1 && work[st->ktarg - 1] < BS_KFAC1 * work[st->ktarg])
                        st->ktarg--;
                    hnew = hopt[st->ktarg];
                }
                break;
            }
        }
        if (st->reject) {
            st->prev_reject = 1;
            st->rej++;
        }
    }
    bs_rhs(f, prm->qm, st->t + h, y, dnew);
    st->rhs++;
    if (!d) break;
    double err_int = bs_prepare_dense(c, h, dnew, ysav, scale, k, y, st->dydx, d);
    double pe = tp_pow(err_int, 1.0 / (2 * k + 3));
    hopt_int = h / (pe > 0.01 ? pe : 0.01);
    if (err_int <= 10.0) break;
    /* interpolation error too large: the step again, with hopt_int */
    hnew = fabs(hopt_int);
    st->reject = 1;
    st->prev_reject = 1;
    st->rej++;
    }
    for (int i = 0; i < BS_N; i++) st->dydx[i] = dnew[i];
    if (d) {
        d->xold = st->t;
        d->h = h;
    }
    st->t += h;
    st->first = 0;
    st->acc++;
    int kopt;
    const int KM1 = c->kmax - 1;
    if (k == 1) {
        kopt = 2;
    } else if (k <= st->ktarg) {
        kopt = k;
        if (work[k - 1] < BS_KFAC1 * work[k]) kopt = k - 1;
        else if (work[k] < BS_KFAC2 * work[k - 1]) kopt = (k + 1 < KM1) ? k + 1 : KM1;
    } else {
        kopt = k - 1;
        if (k > 2 && work[k - 2] < BS_KFAC1 * work[k - 1]) kopt = k - 2;
        if (work[k] < BS_KFAC2 * work[kopt]) kopt = (k < KM1) ? k : KM1;
    }
    if (st->prev_reject) {
        st->ktarg = kopt < k ? kopt : k;
        double a = fabs(h), b = hopt[st->ktarg];
        hnew = a < b ? a : b;
        st->prev_reject = 0;
    } else {
        if (kopt <= k) {
            hnew = hopt[kopt];
        } else {
            if (k < st->ktarg && work[k] < BS_KFAC2 * work[k - 1])
                hnew = hopt[k] * c->cost[kopt + 1] / c->cost[k];
            else
                hnew = hopt[k] * c->cost[kopt] / c->cost[k];
        }
        st->ktarg = kopt;
    }
    if (d) {
        double hi = fabs(hopt_int);
        hnew = hnew < hi ? hnew : hi;
    }
    st->hnext = hnew;
    return 0;
}

/* End time of a particle: t_end, or t + t_end_gyro gyroperiods */
static double bs_tend(const tp_bs_params *prm, const tp_particle *p, double *tg)
{
    double g0 = sqrt(1.0 + p->ux * p->ux + p->uy * p->uy + p->uz * p->uz);
    double T = 2.0 * M_PI * g0 / (fabs(prm->qm) * prm->b_ref);
    *tg = T;
    return prm->t_end_gyro > 0.0 ? p->t + prm->t_end_gyro * T : prm->t_end;
}

int tpref_bs_track(const tpref_field *f, const tp_bs_params *prm, int n, tp_particle *ptl,
                   int *nsteps, long long *totals, int nthreads, double *seconds)
{
    if (n < 0 || prm->kmax < 3 || prm->kmax > 8 || !(prm->rtol > 0.0)) return TP_E_ARG;
    bs_consts c;
    bs_setup(prm->kmax, prm->rtol, 0, &c);
    int cap = prm->max_attempts > 0 ? prm->max_attempts : 10000000;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    long long acc = 0, rej = 0, rhs = 0, failed = 0;
    double t1 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads) reduction(+ : acc, rej, rhs, failed)
    for (int i = 0; i < n; i++) {
        tp_particle *p = &ptl[i];
        double tg;
        double tend = bs_tend(prm, p, &tg);
        double y[BS_N] = {p->x, p->y, p->z, p->ux, p->uy, p->uz};
        bs_state st;
        memset(&st, 0, sizeof(st));
        st.t = p->t;
        st.hnext = prm->h0 > 0.0 ? prm->h0 : 0.01 * tg;
        st.ktarg = c.ktarg0;
        st.first = 1;
        bs_rhs(f, prm->qm, st.t, y, st.dydx);
        st.rhs = 1;
        int fail = 0;
        while (st.t < tend - 1e-13 * fabs(tend)) {
            double h = st.hnext;
            if (st.t + h > tend) h = tend - st.t;
            if (bs_step(f, prm, &c, &st, y, h, cap, NULL)) {
                fail = 1;
                break;
            }
        }
        p->x = y[0]; p->y = y[1]; p->z = y[2];
        p->ux = y[3]; p->uy = y[4]; p->uz = y[5];
        p->t = st.t;
        if (nsteps) nsteps[i] = st.acc;
        acc += st.acc;
        rej += st.rej;
        rhs += st.rhs;
        failed += fail;
    }
    if (seconds) *seconds = omp_get_wtime() - t1;
    if (totals) {
        totals[0] = acc; totals[1] = rej; totals[2] = rhs; totals[3] = failed;
    }
    return TP_OK;
}

/* The adaptive loop with the fixed-step loop's diagnostics: every particle
 * from its time t0 to t0 + estep*dt with dense output, particle_bc after
 * each accepted step (the commented particle_tracking_adaptive,
 * src/tracking.c:518-630), and at the output times of the fixed-step loop
 * (t0 + j*dt for the steps j that record, src/tracking.c:333-372) the
 * spectrum (j % einterval_t == 0) and the D terms (j % dcoeff_interval ==
 * 0) of the state the dense-output polynomial gives there.  Output times
 * strictly inside a step use the polynomial of that step; at the end time
 * the final state is used.  Escapes (bc 1) are binned into the escape
 * spectrum at the frame of the next output time and the particle stops;
 * its later D rows carry stale terms and it counts against the remaining
 * particles as the fixed-step loop counts it.  D terms are kept per
 * particle and row and summed in particle order. */
static void bs_dterms(const tpref_field *f, const tp_particle *p0, const tp_particle *p,
                      double dt, int escaped, int fix_stale, double stale[5], double o[8])
{
    if (escaped) {
        o[0] = 0.0; o[4] = 0.0; o[5] = 0.0;
        if (fix_stale) stale[0] = stale[1] = stale[2] = stale[3] = stale[4] = 0.0;
        o[1] = stale[0]; o[2] = stale[1]; o[3] = stale[2]; o[6] = stale[3]; o[7] = stale[4];
        return;
    }
    tpref_diff_coeff(f, p0, p, dt, o);
    stale[0] = o[1]; stale[1] = o[2]; stale[2] = o[3]; stale[3] = o[6]; stale[4] = o[7];
}

int tpref_bs_track_diag(const tpref_field *f, const tp_bs_params *bs, const tp_track_params *prm,
                        int n, tp_particle *ptl, const tp_particle *ptl_init, int *nsteps,
                        double *espect, double *espect_escape, double *dcoeff, int *nptl_remain,
                        long long *totals, int nthreads, double *seconds)
{
    if (n < 0 || bs->kmax < 3 || bs->kmax > 8 || !(bs->rtol > 0.0) || prm->sstep != 1 ||
        prm->estep < 1 || prm->einterval_t <= 0 || prm->dcoeff_interval <= 0 ||
        prm->nbins < 1 || prm->nt_out < 1 || !(prm->dt > 0.0))
        return TP_E_ARG;
    bs_consts c;
    bs_setup(bs->kmax, bs->rtol, 1, &c);
    const int cap = bs->max_attempts > 0 ? bs->max_attempts : 10000000;
    const int nb = prm->nbins, nt = prm->nt_out, ei = prm->einterval_t, di = prm->dcoeff_interval;
    const int nd = prm->estep / di + 1;
    const int fix_stale = (prm->flags & TP_FLAG_FIX_STALE_DCOEFF) != 0;
    /* frames that exist: steps (f+1)*ei <= estep */
    int nframes = prm->estep / ei;
    if (nframes > nt) nframes = nt;
    double *terms = calloc((size_t)n * nd * 8 + 1, sizeof(double));
    int *out_from = malloc(sizeof(int) * (n + 1));   /* first row counting it out */
    if (!terms || !out_from) { free(terms); free(out_from); return TP_E_NOMEM; }
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    long long acc = 0, rej = 0, rhs = 0, failed = 0;
    unsigned *cnt = calloc((size_t)nthreads * nb * nt * 2 + 1, sizeof(unsigned));
    if (!cnt) { free(terms); free(out_from); return TP_E_NOMEM; }
    double t1 = omp_get_wtime();
#pragma omp parallel num_threads(nthreads) reduction(+ : acc, rej, rhs, failed)
    {
        unsigned *hsp = cnt + (size_t)omp_get_thread_num() * nb * nt * 2;
        unsigned *hes = hsp + (size_t)nb * nt;
        bs_dense *d = malloc(sizeof(bs_dense));
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < n; i++) {
            tp_particle q = ptl[i];
            const tp_particle *p0 = &ptl_init[i];
            double *ti = terms + (size_t)i * nd * 8;
            double stale[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            const double t0 = q.t;
            const double tend = t0 + prm->estep * prm->dt;
            int esc = 0, nf = 0, nr = 1, fail = 0;
            bs_state st;
            memset(&st, 0, sizeof(st));
            tpref_particle_bc(prm, &q, &esc);          /* src/tracking.c:314-332 */
            const int init_esc = esc;
            /* trajectory records (armed by tpref_set_trajectory): the
             * fixed-step loop's rule, the particle at the output steps j =
             * k nsteps_output <= estep, slot 1 on */
            const int toff = (g_traj_nout > 0 && i > 0) ? g_traj_accum[i - 1] : 0;
            const int tcap = g_traj_nout > 0 ? g_traj_accum[i] - toff : 0;
            int tpts = 1, jt = g_traj_nout;
            if (!esc) {
                double y[BS_N] = {q.x, q.y, q.z, q.ux, q.uy, q.uz};
                double g0 = sqrt(1.0 + q.ux * q.ux + q.uy * q.uy + q.uz * q.uz);
                double T = 2.0 * M_PI * g0 / (fabs(bs->qm) * (bs->b_ref > 0.0 ? bs->b_ref : 1.0));
                st.t = t0;
                st.hnext = bs->h0 > 0.0 ? bs->h0 : 0.01 * T;
                st.ktarg = c.ktarg0;
                st.first = 1;
                bs_rhs(f, bs->qm, st.t, y, st.dydx);
                st.rhs = 1;
                while (st.t < tend - 1e-13 * fabs(tend)) {
                    double h = st.hnext;
                    if (st.t + h > tend) h = tend - st.t;
                    if (bs_step(f, bs, &c, &st, y, h, cap, d)) { fail = 1; break; }
                    /* output times strictly inside the step, from the polynomial
                     * (the state relative to the step's start; shifts as then) */
                    tp_particle e = q;
                    for (; nf < nframes; nf++) {
                        const double tf = t0 + (double)((nf + 1) * ei) * prm->dt;
                        if (!(tf < st.t)) break;
                        e.ux = bs_dense_out(d, 3, tf);
                        e.uy = bs_dense_out(d, 4, tf);
                        e.uz = bs_dense_out(d, 5, tf);
                        hsp[nf * nb + tpref_energy_bin(&e, nb, prm->logemin, prm->logde)]++;
                    }
                    for (; nr < nd; nr++) {
                        const double tr = t0 + (double)(nr * di) * prm->dt;
                        if (!(tr < st.t)) break;
                        double yd[BS_N];
                        for (int m = 0; m < BS_N; m++) yd[m] = bs_dense_out(d, m, tr);
                        e.x = yd[0]; e.y = yd[1]; e.z = yd[2];
                        e.ux = yd[3]; e.uy = yd[4]; e.uz = yd[5]; e.t = tr;
                        bs_dterms(f, p0, &e, prm->dt, 0, fix_stale, stale, ti + 8 * nr);
                    }
                    /* trajectory points inside the step: the polynomial's
                     * state, folded into the box as particle_bc folds the
                     * fixed-step loop's records */
                    for (; g_traj_nout > 0 && jt <= prm->estep; jt += g_traj_nout) {
                        const double tj = t0 + (double)jt * prm->dt;
                        if (!(tj < st.t)) break;
                        tp_particle r = q;
                        r.x = bs_dense_out(d, 0, tj); r.y = bs_dense_out(d, 1, tj);
                        r.z = bs_dense_out(d, 2, tj); r.ux = bs_dense_out(d, 3, tj);
                        r.uy = bs_dense_out(d, 4, tj); r.uz = bs_dense_out(d, 5, tj);
                        r.t = tj;
                        if (prm->bc_flag == 0) {
                            int dummy = 0;
                            tpref_particle_bc(prm, &r, &dummy);
                        }
                        if (tpts < tcap) g_traj[toff + tpts] = r;
                        tpts++;
                    }
                    /* the step's end: boundary, escape */
                    q.x = y[0]; q.y = y[1]; q.z = y[2];
                    q.ux = y[3]; q.uy = y[4]; q.uz = y[5]; q.t = st.t;
                    tpref_particle_bc(prm, &q, &esc);
                    y[0] = q.x; y[1] = q.y; y[2] = q.z;
                    if (esc) {
                        if (nf < nt)
                            hes[nf * nb + tpref_energy_bin(&q, nb, prm->logemin, prm->logde)]++;
                        break;
                    }
                }
                if (!esc && !fail) {
                    /* the remaining output times are the end time: the final state */
                    q.t = st.t;
                    for (; nf < nframes; nf++)
                        hsp[nf * nb + tpref_energy_bin(&q, nb, prm->logemin, prm->logde)]++;
                    for (; nr < nd; nr++) {
                        tp_particle e = q;
                        e.t = t0 + (double)(nr * di) * prm->dt;
                        bs_dterms(f, p0, &e, prm->dt, 0, fix_stale, stale, ti + 8 * nr);
                    }
                    for (; g_traj_nout > 0 && jt <= prm->estep; jt += g_traj_nout) {
                        tp_particle r = q;
                        r.t = t0 + (double)jt * prm->dt;
                        if (tpts < tcap) g_traj[toff + tpts] = r;
                        tpts++;
                    }
                }
            }
            if (g_traj_nout > 0 && g_traj_points) g_traj_points[i] = tpts;
            /* rows from nr on: the particle has left (or failed); as the
             * fixed-step loop counts it, an initial escape is out from row
             * 1, an escape before D row r from row r + 1 */
            out_from[i] = !esc ? nd : (init_esc ? 1 : nr + 1);
            for (int r = nr; r < nd; r++) bs_dterms(f, p0, &q, prm->dt, 1, fix_stale, stale, ti + 8 * r);
            ptl[i] = q;
            if (nsteps) nsteps[i] = st.acc;
            acc += st.acc;
            rej += st.rej;
            rhs += st.rhs;
            failed += fail;
        }
        free(d);
    }
    if (seconds) *seconds = omp_get_wtime() - t1;
    /* spectra: integer counts over the threads */
    for (size_t k = 0; k < (size_t)nb * nt; k++) {
        unsigned long long a = 0, b = 0;
        for (int th = 0; th < nthreads; th++) {
            a += cnt[(size_t)th * nb * nt * 2 + k];
            b += cnt[(size_t)th * nb * nt * 2 + (size_t)nb * nt + k];
        }
        if (espect) espect[k] = (double)a;
        if (espect_escape) espect_escape[k] = (double)b;
    }
    /* D sums in particle order; row 0 stays 0 */
    if (dcoeff) {
        memset(dcoeff, 0, sizeof(double) * 8 * nd);
        for (int r = 1; r < nd; r++)
            for (int q = 0; q < 8; q++) {
                double s = 0.0;
                for (int i = 0; i < n; i++) s += terms[((size_t)i * nd + r) * 8 + q];
                dcoeff[(size_t)q * nd + r] = s;
            }
    }
    if (nptl_remain) {
        for (int k = 0; k < nd; k++) {
            int out = 0;
            for (int i = 0; i < n; i++) out += out_from[i] <= k;
            nptl_remain[k] = n - out;
        }
    }
    free(terms);
    free(out_from);
    free(cnt);
    if (totals) {
        totals[0] = acc; totals[1] = rej; totals[2] = rhs; totals[3] = failed;
    }
    return TP_OK;
}

/* ---------------------------------------------------------------------- */
/* Poincare map (run_type 2, src/poincare_map.c:39-258)                    */
/* ---------------------------------------------------------------------- */
/* field_boundary, src/poincare_map.c:326-352: each axis folded once */
static void pmap_fold(const tp_pmap_params *q, double *v)
{
    for (int d = 0; d < 3; d++) {
        const double len = q->hi[d] - q->lo[d];
        if (v[d] > q->hi[d]) v[d] -= len;
    }
    for (int d = 0; d < 3; d++) {
        const double len = q->hi[d] - q->lo[d];
        if (v[d] < q->lo[d]) v[d] += len;
    }
}

/* B / |B| at t = 0 (get_magnetic_field, src/emfields.c:76-96, and the
 * normalisation of src/poincare_map.c:282-285) */
static void pmap_unit(const tpref_field *f, const double *v, double *u)
{
    emf_t e;
    get_emf(f, v[0], v[1], v[2], 0.0, &e);
    double ib = 1.0 / sqrt(e.Bx * e.Bx + e.By * e.By + e.Bz * e.Bz);
    u[0] = e.Bx * ib;
    u[1] = e.By * ib;
    u[2] = e.Bz * ib;
}

/* next_location, src/poincare_map.c:268-320; *folded gets the face
 * coordinate after the initial fold */
static void pmap_rk4(const tpref_field *f, const tp_pmap_params *q, double *p, double *folded)
{
    double k1[3], k2[3], k3[3], k4[3], t[3];
    const double ds = q->step;
    pmap_fold(q, p);
    *folded = p[q->face];
    pmap_unit(f, p, k1);
    for (int d = 0; d < 3; d++) t[d] = p[d] + 0.5 * ds * k1[d];
    pmap_fold(q, t);
    pmap_unit(f, t, k2);
    for (int d = 0; d < 3; d++) t[d] = p[d] + 0.5 * ds * k2[d];
    pmap_fold(q, t);
    pmap_unit(f, t, k3);
    for (int d = 0; d < 3; d++) t[d] = p[d] + ds * k3[d];
    pmap_fold(q, t);
    pmap_unit(f, t, k4);
    for (int d = 0; d < 3; d++) p[d] += (ds / 6) * (k1[d] + 2 * k2[d] + 2 * k3[d] + k4[d]);
}

/* Whether the step from `before` (face coordinate before the step's fold) /
 * `folded` (after it) to `after` crossed the plane c0. */
static int pmap_crossed(const tp_pmap_params *q, double c0, double before, double folded,
                        double after)
{
    if (q->mode == TP_PMAP_REF_COMPAT)
        return !((before - c0) * (after - c0) > 0);     /* src/poincare_map.c:168-171 */
    const double L = q->hi[q->face] - q->lo[q->face];
    double d0 = folded - c0, d1 = after - c0;
    d0 = d0 - L * rint(d0 / L);
    d1 = d1 - L * rint(d1 / L);
    return (d0 * d1 < 0 && fabs(d0 - d1) < 0.5 * L) || (d1 == 0 && d0 != 0);
}

int tpref_poincare_map(const tpref_field *f, const tp_pmap_params *q, int n,
                       const double *starts, double *xpos, double *ypos, double *zpos,
                       int *ncrossed, long long *nsteps, int nthreads, double *seconds)
{
    if (n < 0 || q->face < 0 || q->face > 2 || q->ncross < 0 || !(q->step > 0.0) ||
        !isfinite(q->step) || !isfinite(q->length_max) || !(q->length_max / q->step < 0x1p50))
        return TP_E_ARG;
    if (q->mode != TP_PMAP_REF_COMPAT && q->mode != TP_PMAP_PERIODIC) return TP_E_ARG;
    const int ns = q->ncross + 1;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    double t1 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int i = 0; i < n; i++) {
        double *xs = xpos + (size_t)i * ns, *ys = ypos + (size_t)i * ns,
               *zs = zpos + (size_t)i * ns;
        /* slots a line does not reach keep 2 * hi (src/poincare_map.c:75-80) */
        for (int s = 0; s < ns; s++) {
            xs[s] = q->hi[0] * 2;
            ys[s] = q->hi[1] * 2;
            zs[s] = q->hi[2] * 2;
        }
        double p[3] = {starts[3 * i], starts[3 * i + 1], starts[3 * i + 2]};
        xs[0] = p[0]; ys[0] = p[1]; zs[0] = p[2];
        const double c0 = p[q->face];
        double len = 0.0;
        long long st = 0;
        int k = 1;
        /* src/poincare_map.c:166-184: trace until the next crossing, record
         * it, stop at ncross crossings or when the length reaches the cap */
        while (k <= q->ncross && len < q->length_max) {
            int hit;
            do {
                double before = p[q->face], folded;
                pmap_rk4(f, q, p, &folded);
                len += q->step;
                st++;
                hit = pmap_crossed(q, c0, before, folded, p[q->face]);
            } while (!hit && len < q->length_max);
            if (!(len < q->length_max)) break;
            xs[k] = p[0]; ys[k] = p[1]; zs[k] = p[2];
            k++;
        }
        if (ncrossed) ncrossed[i] = k - 1;
        if (nsteps) nsteps[i] = st;
    }
    if (seconds) *seconds = omp_get_wtime() - t1;
    return TP_OK;
}
