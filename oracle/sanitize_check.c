/*
 * sanitize_check.c -- TEST INFRASTRUCTURE: drives the CPU code of this
 * repository under AddressSanitizer and UndefinedBehaviorSanitizer
 * (`make -C oracle sanitize`, run by tests/test_sanitize_cpu.py).
 *
 * Covered: the oracle (tpref.c, mlpref.c) over every field kind, both
 * pushers, the three boundary conditions, the adaptive stepper with and
 * without dense output, the Poincare map and the surrogate network; the
 * host C of the product (test_particle_amd/csrc/tp_host.c: spectrum files,
 * .gda field input, Poincare start points) and the drop-in's init.dat
 * section reader (test_particle_amd/host/tp_shim_config.c) on the files
 * named on the command line.  Inputs are small and seeded; every result is
 * checked for the invariants that hold whatever the numbers (counts add up,
 * states finite), and the program prints one line per part.  A sanitizer
 * report ends the program with a non-zero status.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include "tpref.h"
#include "../include/tp_host.h"
#include "../test_particle_amd/host/tp_shim_config.h"

int tpref_mlp_nparams(void);
void tpref_mlp_forward(const double *params, const double *q, int n, const double *x,
                       double *out, int emulate);
void tpref_mlp_loss_grad(const double *params, const double *q, int n, const double *x,
                         const double *y, double *grad, double *loss, int emulate);

static int g_fail;

#define CHECK(cond, ...)                                                   \
    do {                                                                   \
        if (!(cond)) {                                                     \
            fprintf(stderr, "sanitize_check: %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);                                  \
            fputc('\n', stderr);                                           \
            g_fail = 1;                                                    \
        }                                                                  \
    } while (0)

static uint64_t g_rng = 0x9E3779B97F4A7C15ull;

static double urand(void)
{
    g_rng = g_rng * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(g_rng >> 11) * 0x1.0p-53;
}

static float *float_grid(int nn, double scale, double bz)
{
    float *g = malloc(sizeof(float) * 3 * (size_t)nn);
    for (int i = 0; i < nn; i++) {
        g[3 * i] = (float)(scale * (urand() - 0.5));
        g[3 * i + 1] = (float)(scale * (urand() - 0.5));
        g[3 * i + 2] = (float)(bz + scale * (urand() - 0.5));
    }
    return g;
}

static void particles(tp_particle *p, int n, const double lo[3], const double hi[3], double u)
{
    for (int i = 0; i < n; i++) {
        p[i].x = lo[0] + (hi[0] - lo[0]) * urand();
        p[i].y = lo[1] + (hi[1] - lo[1]) * urand();
        p[i].z = lo[2] + (hi[2] - lo[2]) * urand();
        p[i].qm = 1.0;
        p[i].ux = u * (2.0 * urand() - 1.0);
        p[i].uy = u * (2.0 * urand() - 1.0);
        p[i].uz = u * (2.0 * urand() - 1.0);
        p[i].t = 0.0;
        p[i].xshift = p[i].yshift = p[i].zshift = 0.0;
    }
}

static tp_track_params track_params(int pusher, int bc, const double lo[3], const double hi[3],
                                     double dt, int estep)
{
    tp_track_params p;
    memset(&p, 0, sizeof(p));
    p.pusher = pusher;
    p.precision = TP_FP64;
    p.bc_flag = bc;
    p.nbins = 24;
    p.nt_out = 3;
    p.sstep = 1;
    p.estep = estep;
    p.einterval_t = estep / 3;
    p.dcoeff_interval = 50;
    p.dt = dt;
    p.qm = 1.0;
    p.xmin = lo[0]; p.xmax = hi[0];
    p.ymin = lo[1]; p.ymax = hi[1];
    p.zmin = lo[2]; p.zmax = hi[2];
    p.logemin = log10(1e-4);
    p.logde = (log10(10.0) - log10(1e-4)) / (p.nbins - 2);
    return p;
}

/* a fixed-step run of both pushers and the three boundary conditions */
static void check_fixed(const char *name, const tpref_field *f, const double lo[3],
                        const double hi[3], const double plo[3], const double phi[3],
                        double dt)
{
    enum { N = 29, STEPS = 150 };
    const int nd = STEPS / 50 + 1;
    tp_particle p0[N], p[N];
    int nsteps[N], remain[8];
    double esp[3 * 24], ese[3 * 24], dco[8 * 8];
    particles(p0, N, plo, phi, 0.4);
    long total = 0;
    for (int pusher = 0; pusher < 2; pusher++)
        for (int bc = 0; bc < 3; bc++) {
            tp_track_params prm = track_params(pusher, bc, lo, hi, dt, STEPS);
            memcpy(p, p0, sizeof(p));
            memset(nsteps, 0, sizeof(nsteps));
            int rc = tpref_track_fixed(f, &prm, N, p, p0, nsteps, esp, ese, dco, remain, 2,
                                       NULL);
            CHECK(rc == 0, "%s: track_fixed pusher %d bc %d returned %d", name, pusher, bc, rc);
            for (int t = 0; t < 3; t++) {
                double c = 0.0;
                for (int b = 0; b < 24; b++) c += esp[t * 24 + b] + ese[t * 24 + b];
                CHECK(c <= N && (bc == 1 || c == N), "%s: frame %d holds %g particles", name, t,
                      c);
            }
            for (int k = 0; k < nd; k++) CHECK(remain[k] >= 0 && remain[k] <= N, "%s", name);
            for (int i = 0; i < N; i++) {
                /* the count starts at 1 (src/tracking.c:128) */
                CHECK(nsteps[i] >= 1 && nsteps[i] <= STEPS + 1, "%s: nsteps %d", name,
                      nsteps[i]);
                total += nsteps[i];
            }
        }
    printf("fixed %-10s %ld particle-steps\n", name, total);
}

static void check_eval(const char *name, const tpref_field *f, const double lo[3],
                       const double hi[3])
{
    enum { N = 64 };
    double pts[4 * N], out[6 * N];
    for (int i = 0; i < N; i++) {
        /* inside the box, on its faces and one cell beyond (wrapped) */
        for (int k = 0; k < 3; k++) {
            const double w = hi[k] - lo[k];
            pts[4 * i + k] = i % 4 == 0 ? lo[k] : i % 4 == 1 ? hi[k]
                           : i % 4 == 2 ? lo[k] + w * urand() : hi[k] + 0.1 * w * urand();
        }
        pts[4 * i + 3] = urand();
    }
    tpref_eval_field(f, N, pts, out);
    int finite = 1;
    for (int i = 0; i < 6 * N; i++) finite &= isfinite(out[i]) != 0;
    CHECK(finite, "%s: non-finite field value", name);
    printf("eval  %-10s ok\n", name);
}

static void check_adaptive(const char *name, const tpref_field *f, const double lo[3],
                           const double hi[3])
{
    enum { N = 9 };
    tp_particle p0[N], p[N];
    int nsteps[N], remain[8];
    long long tot[4];
    double esp[3 * 24], ese[3 * 24], dco[8 * 8];
    particles(p0, N, lo, hi, 0.3);
    tp_bs_params b;
    memset(&b, 0, sizeof(b));
    b.rtol = 1e-8;
    b.atol = 1e-10;
    b.qm = 1.0;
    b.t_end_gyro = 0.3;
    b.b_ref = 1.0;
    b.kmax = 8;
    memcpy(p, p0, sizeof(p));
    int rc = tpref_bs_track(f, &b, N, p, nsteps, tot, 1, NULL);
    CHECK(rc == 0 && tot[3] == 0, "%s: bs_track rc %d failed %lld", name, rc, tot[3]);
    for (int bc = 0; bc < 2; bc++) {
        tp_track_params prm = track_params(TP_PUSH_WIRZ, bc, lo, hi, 0.05, 100);
        memcpy(p, p0, sizeof(p));
        rc = tpref_bs_track_diag(f, &b, &prm, N, p, p0, nsteps, esp, ese, dco, remain, tot, 1,
                                 NULL);
        CHECK(rc == 0, "%s: bs_track_diag bc %d rc %d", name, bc, rc);
        double c = 0.0;
        for (int k = 0; k < 24; k++) c += esp[2 * 24 + k] + ese[2 * 24 + k];
        CHECK(c <= N && (bc == 1 || c == N), "%s: last frame holds %g", name, c);
    }
    printf("bs    %-10s %lld accepted\n", name, tot[0]);
}

static void check_pmap(const char *name, const tpref_field *f, const double lo[3],
                       const double hi[3])
{
    enum { NH = 4, NV = 3, NC = 3 };
    const double p0[3] = {lo[0], lo[1], 0.5 * (lo[2] + hi[2])};
    double starts[3 * NH * NV];
    int np = tp_host_pmap_starts(2, 0, p0, lo[0] + 0.1, hi[0] - 0.1, lo[1] + 0.1, hi[1] - 0.1,
                                 NH, NV, starts);
    CHECK(np == NH * NV, "%s: pmap_starts gave %d", name, np);
    if (np != NH * NV) return;
    double xs[NH * NV * (NC + 1)], ys[NH * NV * (NC + 1)], zs[NH * NV * (NC + 1)];
    int ncr[NH * NV];
    long long ns[NH * NV];
    tp_pmap_params prm;
    memset(&prm, 0, sizeof(prm));
    prm.face = 2;
    prm.ncross = NC;
    prm.mode = TP_PMAP_REF_COMPAT;
    prm.step = 0.05;
    prm.length_max = 40.0;
    for (int k = 0; k < 3; k++) {
        prm.lo[k] = lo[k];
        prm.hi[k] = hi[k];
    }
    int rc = tpref_poincare_map(f, &prm, np, starts, xs, ys, zs, ncr, ns, 2, NULL);
    CHECK(rc == 0, "%s: poincare_map rc %d", name, rc);
    long long steps = 0;
    for (int i = 0; i < np; i++) {
        CHECK(ncr[i] >= 0 && ncr[i] <= NC, "%s: %d crossings", name, ncr[i]);
        steps += ns[i];
    }
    printf("pmap  %-10s %lld steps\n", name, steps);
}

static void check_mlp(void)
{
    const int np = tpref_mlp_nparams();
    enum { N = 7 };
    double *w = malloc(sizeof(double) * (size_t)np), *g = malloc(sizeof(double) * (size_t)np);
    double q[3 * 128], x[3 * N], y[3 * N], out[3 * N], loss = 0.0;
    for (int i = 0; i < np; i++) w[i] = 0.1 * (urand() - 0.5);
    for (int i = 0; i < 3 * 128; i++) q[i] = floor(8.0 * urand()) - 4.0;
    for (int i = 0; i < 3 * N; i++) {
        x[i] = urand();
        y[i] = urand() - 0.5;
    }
    for (int em = 0; em < 2; em++) {
        tpref_mlp_forward(w, q, N, x, out, em);
        tpref_mlp_loss_grad(w, q, N, x, y, g, &loss, em);
        int finite = isfinite(loss) != 0;
        for (int i = 0; i < 3 * N; i++) finite &= isfinite(out[i]) != 0;
        for (int i = 0; i < np; i++) finite &= isfinite(g[i]) != 0;
        CHECK(finite, "mlp: non-finite output (emulate %d)", em);
    }
    printf("mlp   %d parameters, loss %.3e\n", np, loss);
    free(w);
    free(g);
}

/* the host C of the product: spectrum files and .gda field input */
static void check_host(const char *dir)
{
    enum { NB = 12, NT = 2, ND = 3, NX = 5, NY = 4, NZ = 3 };
    char path[4096], path2[4096];
    double ebins[NB], ewidth[NB], counts[NB], rows[NT * NB], d[8 * ND];
    int remain[ND] = {5, 4, 3};
    tp_particle p[40];
    const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    particles(p, 40, lo, hi, 0.5);
    int rc = tp_host_spectrum_bins(NB, 1e-4, 10.0, ebins, ewidth);
    CHECK(rc == 0, "spectrum_bins rc %d", rc);
    rc = tp_host_spectrum_t0(40, p, NB, 1e-4, 10.0, counts);
    double c = 0.0;
    for (int k = 0; k < NB; k++) c += counts[k];
    CHECK(rc == 0 && c == 40.0, "spectrum_t0 rc %d holds %g", rc, c);
    for (int k = 0; k < NT * NB; k++) rows[k] = (double)(k % 5);
    for (int k = 0; k < 8 * ND; k++) d[k] = urand();
    snprintf(path, sizeof(path), "%s/espectrum.dat", dir);
    snprintf(path2, sizeof(path2), "%s/espect_escape.dat", dir);
    rc = tp_host_write_espectrum_t0(path, path2, NB, 1e-4, 10.0, counts, 1);
    CHECK(rc == 0, "write_espectrum_t0 rc %d", rc);
    rc = tp_host_append_espectrum(path, NB, NT, 1e-4, 10.0, rows);
    CHECK(rc == 0, "append_espectrum rc %d", rc);
    snprintf(path, sizeof(path), "%s/diff_coeffs.dat", dir);
    rc = tp_host_write_diff_coeffs(path, ND, d, remain);
    CHECK(rc == 0, "write_diff_coeffs rc %d", rc);

    /* three slices of the six component files, then a short file */
    static const char *const comp[6] = {"ex", "ey", "ez", "bx", "by", "bz"};
    const int nn = NX * NY * NZ;
    float buf[NX * NY * NZ];
    for (int ct = 0; ct < 3; ct++)
        for (int k = 0; k < 6; k++) {
            snprintf(path, sizeof(path), "%s/%s_%d.gda", dir, comp[k], ct);
            FILE *fp = fopen(path, "wb");
            CHECK(fp != NULL, "cannot write %s", path);
            if (!fp) return;
            for (int i = 0; i < nn; i++) buf[i] = (float)(100 * ct + 10 * k) + (float)i * 0.01f;
            fwrite(buf, sizeof(float), (size_t)(k == 5 && ct == 2 ? nn - 1 : nn), fp);
            fclose(fp);
        }
    snprintf(path, sizeof(path), "%s/", dir);
    float *en = malloc(sizeof(float) * 3 * nn), *bn = malloc(sizeof(float) * 3 * nn);
    float *ep = malloc(sizeof(float) * 3 * nn), *bp = malloc(sizeof(float) * 3 * nn);
    rc = tp_host_read_pic_gda(path, 1, 3, 0, NX, NY, NZ, 2.0, 3.0, en, bn, NULL, NULL);
    CHECK(rc == 0 && en[0] == 200.0f && bn[3 * 7 + 2] == (float)(3.0 * (double)(150.0f +
          7 * 0.01f)), "read_pic_gda (one slice) rc %d", rc);
    rc = tp_host_read_pic_gda(path, 1, 2, 1, NX, NY, NZ, 1.0, 1.0, en, bn, ep, bp);
    CHECK(rc == 0 && en[0] == 100.0f && ep[0] == 0.0f, "read_pic_gda (two slices) rc %d", rc);
    rc = tp_host_read_pic_gda(path, 1, 3, 1, NX, NY, NZ, 1.0, 1.0, en, bn, ep, bp);
    CHECK(rc == TP_E_ARG, "read_pic_gda of a short file gave %d", rc);
    rc = tp_host_read_pic_gda(path, 7, 9, 0, NX, NY, NZ, 1.0, 1.0, en, bn, NULL, NULL);
    CHECK(rc == TP_E_ARG, "read_pic_gda of a missing file gave %d", rc);
    free(en); free(bn); free(ep); free(bp);
    printf("host  spectrum files and .gda input ok\n");
}

static void check_configs(int argc, char **argv)
{
    for (int i = 1; i < argc; i++) {
        tp_shim_config c;
        char err[200];
        int rc = tp_shim_config_read(argv[i], &c, err, sizeof(err));
        if (rc == 0)
            printf("config %s: found %d pusher %d precision %d flags %d rtol %.17g atol %.17g "
                   "h0 %.17g kmax %d\n", argv[i], c.found, c.pusher, c.precision, c.flags,
                   c.bs_rtol, c.bs_atol, c.bs_h0, c.bs_kmax);
        else
            printf("config %s: error: %s\n", argv[i], err);
    }
}

int main(int argc, char **argv)
{
    const double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {6.0, 5.0, 4.0};
    const double in_lo[3] = {1.0, 1.0, 1.0}, in_hi[3] = {5.0, 4.0, 3.0};
    const int nx = 7, ny = 6, nz = 5, nn = nx * ny * nz;
    tpref_field f;

    memset(&f, 0, sizeof(f));
    f.kind = TPREF_FIELD_TEST;
    check_eval("test", &f, lo, hi);
    check_fixed("test", &f, lo, hi, in_lo, in_hi, 0.05);
    check_adaptive("test", &f, lo, hi);

    float *e = float_grid(nn, 0.2, 0.0), *b = float_grid(nn, 1.0, 2.0);
    float *e2 = float_grid(nn, 0.2, 0.0), *b2 = float_grid(nn, 1.0, 2.0);
    for (int mode = 0; mode < 2; mode++) {
        memset(&f, 0, sizeof(f));
        f.kind = TPREF_FIELD_GRID;
        f.nx = nx; f.ny = ny; f.nz = nz;
        memcpy(f.lo, lo, sizeof(f.lo));
        memcpy(f.hi, hi, sizeof(f.hi));
        f.egrid = e;
        f.bgrid = b;
        f.gather_mode = mode;
        check_eval(mode ? "grid-tb" : "grid", &f, lo, hi);
        check_fixed(mode ? "grid-tb" : "grid", &f, lo, hi, lo, hi, 0.05);
        if (mode == 0) {
            check_adaptive("grid", &f, lo, hi);
            check_pmap("grid", &f, lo, hi);
            f.egrid_b = e2;
            f.bgrid_b = b2;
            f.grid_dt = 1.7;
            check_eval("grid2", &f, lo, hi);
            check_fixed("grid2", &f, lo, hi, lo, hi, 0.05);
        }
    }
    memset(&f, 0, sizeof(f));
    f.kind = TPREF_FIELD_MHD;
    f.nx = nx; f.ny = ny; f.nz = nz;
    memcpy(f.lo, lo, sizeof(f.lo));
    memcpy(f.hi, hi, sizeof(f.hi));
    f.egrid = e;
    f.bgrid = b;
    check_eval("mhd", &f, lo, hi);
    check_fixed("mhd", &f, lo, hi, lo, hi, 0.05);

    double *ed = malloc(sizeof(double) * 3 * nn), *bd = malloc(sizeof(double) * 3 * nn);
    for (int i = 0; i < 3 * nn; i++) {
        ed[i] = e[i];
        bd[i] = b[i];
    }
    memset(&f, 0, sizeof(f));
    f.kind = TPREF_FIELD_GRIDD;
    f.nx = nx; f.ny = ny; f.nz = nz;
    memcpy(f.lo, lo, sizeof(f.lo));
    memcpy(f.hi, hi, sizeof(f.hi));
    f.egrid_d = ed;
    f.bgrid_d = bd;
    check_eval("grid-f64", &f, lo, hi);
    check_fixed("grid-f64", &f, lo, hi, lo, hi, 0.05);

    /* ABC with and without the velocity grid */
    const double two_pi = 6.283185307179586;
    const double abc_hi[3] = {two_pi, two_pi, two_pi};
    double *vg = malloc(sizeof(double) * 3 * nn);
    for (int i = 0; i < 3 * nn; i++) vg[i] = 0.1 * (urand() - 0.5);
    for (int withv = 0; withv < 2; withv++) {
        memset(&f, 0, sizeof(f));
        f.kind = TPREF_FIELD_ABC;
        f.abc[0] = 1.0; f.abc[1] = sqrt(2.0 / 3.0); f.abc[2] = sqrt(1.0 / 3.0);
        f.abc[3] = 1.0; f.abc[4] = 1.0;
        if (withv) {
            f.vnx = nx; f.vny = ny; f.vnz = nz;
            memcpy(f.vhi, abc_hi, sizeof(f.vhi));
            f.vgrid = vg;
        }
        check_eval(withv ? "abc-v" : "abc", &f, lo, abc_hi);
        check_fixed(withv ? "abc-v" : "abc", &f, lo, abc_hi, lo, abc_hi, 0.05);
    }

    /* multi-mode turbulence */
    enum { NM = 16 };
    double modes[7 * NM];
    for (int m = 0; m < NM; m++) {
        for (int k = 0; k < 3; k++) modes[7 * m + k] = floor(6.0 * urand()) - 3.0;
        for (int k = 3; k < 6; k++) modes[7 * m + k] = 0.2 * (urand() - 0.5);
        modes[7 * m + 6] = two_pi * urand();
    }
    memset(&f, 0, sizeof(f));
    f.kind = TPREF_FIELD_MODES;
    f.nmodes = NM;
    f.modes = modes;
    f.b0z = 1.0;
    check_eval("modes", &f, lo, hi);
    check_fixed("modes", &f, lo, hi, lo, hi, 0.05);
    check_adaptive("modes", &f, lo, hi);

    /* WLCS: a wire and a loop per row (14 values, src/wlcs.c:166-212) */
    const double wl[3 * 14] = {
        1.0, 0.5, 0.5, 0.5, 0.0, 0.0, 0.0, 0.0, 1.0, 0.5, 0.5, 0.5, 0.0, 0.0,
        0.7, 0.2, 0.9, 0.4, 0.3, 1.1, 10.0, 0.6, 0.3, 0.8, 0.1, 0.6, 0.5, 2.0,
        -0.4, 0.8, 0.1, 0.7, 2.0, 4.0, 5.0, 1.2, 0.7, 0.3, 0.6, 0.2, 1.3, 0.4};
    const double wl_lo[3] = {-1.0, -1.0, -1.0}, wl_hi[3] = {2.0, 2.0, 2.0};
    for (int ib = 0; ib < 2; ib++) {
        memset(&f, 0, sizeof(f));
        f.kind = TPREF_FIELD_WLCS;
        f.nwlcs = 3;
        f.wlcs_rows = wl;
        f.ibessel = ib;
        f.omega = 0.001;
        f.norm_i = 1.0;
        check_eval(ib ? "wlcs-j" : "wlcs", &f, wl_lo, wl_hi);
        check_fixed(ib ? "wlcs-j" : "wlcs", &f, wl_lo, wl_hi, wl_lo, wl_hi, 0.02);
    }

    check_mlp();

    char dir[] = "/tmp/tp_sanitize_XXXXXX";
    if (mkdtemp(dir)) {
        check_host(dir);
        static const char *const names[] = {"espectrum.dat", "espect_escape.dat",
                                            "diff_coeffs.dat"};
        char path[4096];
        for (int k = 0; k < 3; k++) {
            snprintf(path, sizeof(path), "%s/%s", dir, names[k]);
            unlink(path);
        }
        static const char *const comp[6] = {"ex", "ey", "ez", "bx", "by", "bz"};
        for (int ct = 0; ct < 3; ct++)
            for (int k = 0; k < 6; k++) {
                snprintf(path, sizeof(path), "%s/%s_%d.gda", dir, comp[k], ct);
                unlink(path);
            }
        rmdir(dir);
    } else {
        CHECK(0, "mkdtemp failed");
    }

    check_configs(argc, argv);

    free(e); free(b); free(e2); free(b2); free(ed); free(bd); free(vg);
    printf("%s\n", g_fail ? "sanitize_check: FAILED" : "sanitize_check: ok");
    return g_fail;
}
