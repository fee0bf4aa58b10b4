/*
 * tp_host.c -- host-side output writers (include/tp_host.h).
 *
 * The reference writes its spectra and diffusion coefficients from the
 * MPI-reduced arrays in src/diagnostics.c; these functions produce the same
 * files from the raw arrays the GPU library returns, with the same formulas
 * for the bin centres and widths and the same printf formats.
 */
#define _FILE_OFFSET_BITS 64
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/types.h>

#include "../../include/tp_host.h"

int tp_host_spectrum_bins(int nbins, double emin, double emax, double *ebins,
                          double *einterval)
{
    if (nbins < 3 || !(emin > 0.0) || !(emax > emin)) return TP_E_ARG;
    double logemin = log10(emin), logemax = log10(emax);
    double logde = (logemax - logemin) / (nbins - 2);
    if (ebins) {
        ebins[0] = 0.5 * emin;
        for (int i = 1; i < nbins - 1; i++) ebins[i] = pow(10, logemin + logde * (i - 0.5));
        ebins[nbins - 1] = pow(10, logemax + 0.5 * logde);
    }
    if (einterval) {
        einterval[0] = emin;
        for (int i = 1; i < nbins - 1; i++)
            einterval[i] = pow(10, logemin) * (pow(10, logde * i) - pow(10, logde * (i - 1)));
        einterval[nbins - 1] = pow(10, logemin) * (pow(10, logde * nbins) -
                                                   pow(10, logde * (nbins - 1)));
    }
    return TP_OK;
}

int tp_host_spectrum_t0(int n, const tp_particle *ptl, int nbins, double emin,
                        double emax, double *counts)
{
    if (n < 0 || (n > 0 && !ptl) || !counts || nbins < 3 || !(emin > 0.0) || !(emax > emin))
        return TP_E_ARG;
    double logemin = log10(emin);
    double logde = (log10(emax) - logemin) / (nbins - 2);
    for (int i = 0; i < nbins; i++) counts[i] = 0.0;
    for (int i = 0; i < n; i++) {
        const tp_particle *p = &ptl[i];
        double gama = sqrt(1.0 + p->ux * p->ux + p->uy * p->uy + p->uz * p->uz);
        double v = (log10(gama - 1.0) - logemin) / logde;
        /* int conversion truncates; out-of-range and NaN give INT_MIN on
         * x86-64, hence bin 0 */
        int ibin = (v > -2147483649.0 && v < 2147483648.0) ? (int)v : 0;
        if (ibin < 0) ibin = 0;
        else if (ibin > nbins - 1) ibin = nbins - 1;
        counts[ibin] += 1.0;
    }
    return TP_OK;
}

int tp_host_write_espectrum_t0(const char *fname, const char *fname_escape, int nbins,
                               double emin, double emax, const double *counts_t0,
                               int bc_flag)
{
    if (!fname || !counts_t0 || nbins < 3) return TP_E_ARG;
    double *eb = malloc(sizeof(double) * nbins), *ei = malloc(sizeof(double) * nbins);
    if (!eb || !ei || tp_host_spectrum_bins(nbins, emin, emax, eb, ei) != TP_OK) {
        free(eb); free(ei);
        return TP_E_ARG;
    }
    int rc = TP_OK;
    FILE *fp = fopen(fname, "w");
    if (!fp) rc = TP_E_ARG;
    if (fp) {
        for (int i = 0; i < nbins; i++) fprintf(fp, "%10.4e ", eb[i]);
        fprintf(fp, "\n");
        for (int i = 0; i < nbins; i++) fprintf(fp, "%10.4e ", counts_t0[i] / ei[i]);
        fprintf(fp, "\n");
        fclose(fp);
    }
    if (rc == TP_OK && bc_flag == 1 && fname_escape) {
        fp = fopen(fname_escape, "w");
        if (!fp) rc = TP_E_ARG;
        if (fp) {
            for (int i = 0; i < nbins; i++) fprintf(fp, "%10.4e ", eb[i]);
            fprintf(fp, "\n");
            for (int i = 0; i < nbins; i++) fprintf(fp, "%10.4e ", 0.0);
            fprintf(fp, "\n");
            fclose(fp);
        }
    }
    free(eb);
    free(ei);
    return rc;
}

int tp_host_append_espectrum(const char *fname, int nbins, int nt_out, double emin,
                             double emax, const double *counts)
{
    if (!fname || !counts || nbins < 3 || nt_out < 0) return TP_E_ARG;
    double *ei = malloc(sizeof(double) * nbins);
    if (!ei || tp_host_spectrum_bins(nbins, emin, emax, NULL, ei) != TP_OK) {
        free(ei);
        return TP_E_ARG;
    }
    FILE *fp = fopen(fname, "a");
    if (!fp) {
        free(ei);
        return TP_E_ARG;
    }
    for (int i = 0; i < nt_out; i++) {
        for (int j = 0; j < nbins; j++) fprintf(fp, "%10.4e ", counts[(size_t)i * nbins + j] / ei[j]);
        fprintf(fp, "\n");
    }
    fclose(fp);
    free(ei);
    return TP_OK;
}

int tp_host_write_diff_coeffs(const char *fname, int nd, const double *d, const int *remain)
{
    if (!fname || nd < 0 || (nd > 0 && (!d || !remain))) return TP_E_ARG;
    FILE *fp = fopen(fname, "w");
    if (!fp) return TP_E_ARG;
    fprintf(fp, "drr        dxx        dyy        dzz        dpp");
    fprintf(fp, "        duu        daa        dee        nptl\n");
    for (int i = 0; i < nd; i++) {
        fprintf(fp, "%10.4e %10.4e %10.4e %10.4e %10.4e %10.4e %10.4e %10.4e %d\n",
                d[i], d[nd + i], d[2 * nd + i], d[3 * nd + i], d[4 * nd + i],
                d[5 * nd + i], d[6 * nd + i], d[7 * nd + i], remain[i]);
    }
    fclose(fp);
    return TP_OK;
}

/* one component file into every third float of out, scaled (float)(v * scale) */
static int read_component(const char *path, const char *comp, int ct, size_t nn,
                          double scale, float *buf, float *out, int k)
{
    char fname[4096];
    int m = snprintf(fname, sizeof(fname), "%s%s_%d.gda", path, comp, ct);
    if (m < 0 || (size_t)m >= sizeof(fname)) return TP_E_ARG;
    FILE *fp = fopen(fname, "rb");
    if (!fp) return TP_E_ARG;
    size_t got = fread(buf, sizeof(float), nn, fp);
    fclose(fp);
    if (got != nn) return TP_E_ARG;
    for (size_t i = 0; i < nn; i++) out[3 * i + k] = (float)((double)buf[i] * scale);
    return TP_OK;
}

static int read_slice(const char *path, int ct, size_t nn, double e0, double b0, float *buf,
                      float *e, float *b)
{
    static const char *ec[3] = {"ex", "ey", "ez"}, *bc[3] = {"bx", "by", "bz"};
    for (int k = 0; k < 3; k++) {
        int rc = read_component(path, ec[k], ct, nn, e0, buf, e, k);
        if (rc) return rc;
        rc = read_component(path, bc[k], ct, nn, b0, buf, b, k);
        if (rc) return rc;
    }
    return TP_OK;
}

int tp_host_read_pic_gda(const char *path, int ct, int nt, int two_slices, int nx, int ny,
                         int nz, double e0, double b0, float *e_next, float *b_next,
                         float *e_prev, float *b_prev)
{
    if (!path || nx < 1 || ny < 1 || nz < 1 || ct < 0 || !e_next || !b_next) return TP_E_ARG;
    if (two_slices && (nt < 1 || ct >= nt || !e_prev || !b_prev)) return TP_E_ARG;
    const size_t nn = (size_t)nx * ny * nz;
    float *buf = (float *)malloc(nn * sizeof(float));
    if (!buf) return TP_E_NOMEM;
    int rc;
    if (!two_slices) {
        rc = read_slice(path, ct, nn, e0, b0, buf, e_next, b_next);
    } else {
        /* src/magnetic_field.c:470-524: previous = ct - 1 (ct at 0),
         * next = ct + 1 (ct at the last slice) */
        rc = read_slice(path, ct > 0 ? ct - 1 : ct, nn, e0, b0, buf, e_prev, b_prev);
        if (!rc) rc = read_slice(path, ct < nt - 1 ? ct + 1 : ct, nn, e0, b0, buf, e_next, b_next);
    }
    free(buf);
    return rc;
}

/* the open component files of the slices being loaded, read in slabs */
typedef struct {
    FILE *fp[2][6];
    size_t inner;
} gda_files;

static int gda_read(void *user, int slice, int comp, long long s0, long long s1, float *dst)
{
    gda_files *g = (gda_files *)user;
    FILE *fp = g->fp[slice][comp];
    const size_t cnt = (size_t)(s1 - s0) * g->inner;
    if (fseeko(fp, (off_t)s0 * (off_t)g->inner * (off_t)sizeof(float), SEEK_SET) != 0) return 1;
    return fread(dst, sizeof(float), cnt, fp) == cnt ? 0 : 1;
}

int tp_host_load_pic_gda(tp_ctx *ctx, const char *path, int ct, int nt, int two_slices, int nx,
                         int ny, int nz, const double lo[3], const double hi[3], double e0,
                         double b0, double grid_dt, int gather_mode)
{
    static const char *names[6] = {"ex", "ey", "ez", "bx", "by", "bz"};
    if (!ctx || !path || nx < 1 || ny < 1 || nz < 1 || ct < 0) return TP_E_ARG;
    if (two_slices != 0 && two_slices != 1) return TP_E_ARG;
    if (two_slices && (nt < 1 || ct >= nt)) return TP_E_ARG;
    /* slice 0 = next (ct + 1, ct at the last slice), slice 1 = previous
     * (ct - 1, ct at 0), src/magnetic_field.c:470-524; one slice: ct */
    const int cts[2] = {two_slices ? (ct < nt - 1 ? ct + 1 : ct) : ct, ct > 0 ? ct - 1 : ct};
    gda_files g;
    memset(&g, 0, sizeof(g));
    g.inner = (size_t)nx * ny;
    const long long need = (long long)nx * ny * nz * (long long)sizeof(float);
    int rc = TP_OK;
    for (int sl = 0; sl <= two_slices && rc == TP_OK; sl++)
        for (int k = 0; k < 6 && rc == TP_OK; k++) {
            char fname[4096];
            int m = snprintf(fname, sizeof(fname), "%s%s_%d.gda", path, names[k], cts[sl]);
            if (m < 0 || (size_t)m >= sizeof(fname)) { rc = TP_E_ARG; break; }
            g.fp[sl][k] = fopen(fname, "rb");
            if (!g.fp[sl][k] || fseeko(g.fp[sl][k], 0, SEEK_END) != 0 ||
                ftello(g.fp[sl][k]) < need)
                rc = TP_E_ARG;
        }
    if (rc == TP_OK)
        rc = tp_gpu_load_field_grid(ctx, nx, ny, nz, lo, hi, TP_PLANES_X_FASTEST, two_slices, e0,
                                    b0, grid_dt, gather_mode, gda_read, &g, 0);
    for (int sl = 0; sl < 2; sl++)
        for (int k = 0; k < 6; k++)
            if (g.fp[sl][k]) fclose(g.fp[sl][k]);
    return rc;
}

int tp_host_pmap_starts(int face, int single, const double p0[3], double min_h, double max_h,
                        double min_v, double max_v, int nh, int nv, double *starts)
{
    if (face < 0 || face > 2 || !p0 || !starts) return TP_E_ARG;
    if (single) {
        starts[0] = p0[0]; starts[1] = p0[1]; starts[2] = p0[2];
        return 1;
    }
    if (nh < 1 || nv < 1 || (long long)nh * nv > 0x7fffffffLL) return TP_E_ARG;
    const double dh = nh > 1 ? (max_h - min_h) / (nh - 1) : max_h - min_h;
    const double dv = nv > 1 ? (max_v - min_v) / (nv - 1) : max_v - min_v;
    /* the two in-plane axes in increasing order */
    const int ah = face == 0 ? 1 : 0;
    const int av = face == 2 ? 1 : 2;
    const int np = nh * nv;
    for (int ip = 0; ip < np; ip++) {
        double *s = starts + 3 * (size_t)ip;
        s[0] = p0[0]; s[1] = p0[1]; s[2] = p0[2];
        s[ah] = dh * (ip % nh) + min_h;
        s[av] = dv * (ip / nh) + min_v;
    }
    return np;
}
