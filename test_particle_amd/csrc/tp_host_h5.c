/*
 * tp_host_h5.c -- the reference's PIC HDF5 field input (libtph5.so, the
 * functions of include/tp_h5.h).  Plain C over the serial HDF5 library; the
 * GPU load goes through libtpgpu.so's tp_gpu_load_field_grid.
 */
#include <stdlib.h>
#include <string.h>
#include <hdf5.h>
#include "tp_gpu.h"
#include "tp_h5.h"

static const char *const tp_h5_dsets[6] = {"ex", "ey", "ez", "cbx", "cby", "cbz"};

/* the leading nx x ny x nz block of one float dataset (read_data_serial_float_h5,
 * src/data_io.c:52-75) */
static int read_block(hid_t group, const char *name, int nx, int ny, int nz, float *out)
{
    hid_t d = H5Dopen2(group, name, H5P_DEFAULT);
    if (d < 0) return TP_E_ARG;
    int rc = TP_OK;
    hid_t fs = H5Dget_space(d);
    hsize_t dims[3] = {0, 0, 0};
    if (fs < 0 || H5Sget_simple_extent_ndims(fs) != 3) {
        rc = TP_E_ARG;
    } else {
        H5Sget_simple_extent_dims(fs, dims, NULL);
        if (dims[0] < (hsize_t)nx || dims[1] < (hsize_t)ny || dims[2] < (hsize_t)nz) rc = TP_E_ARG;
    }
    if (rc == TP_OK) {
        hsize_t count[3] = {(hsize_t)nx, (hsize_t)ny, (hsize_t)nz};
        hsize_t offset[3] = {0, 0, 0};
        hid_t ms = H5Screate_simple(3, count, NULL);
        if (ms < 0 || H5Sselect_hyperslab(fs, H5S_SELECT_SET, offset, NULL, count, NULL) < 0 ||
            H5Dread(d, H5T_NATIVE_FLOAT, ms, fs, H5P_DEFAULT, out) < 0)
            rc = TP_E_ARG;
        if (ms >= 0) H5Sclose(ms);
    }
    if (fs >= 0) H5Sclose(fs);
    H5Dclose(d);
    return rc;
}

int tp_h5_read_pic_planes(const char *fname, const char *gname, int nx, int ny, int nz,
                          float *planes)
{
    if (!fname || !gname || !planes || nx < 1 || ny < 1 || nz < 1) return TP_E_ARG;
    const size_t nn = (size_t)nx * ny * nz;
    H5E_auto2_t fn;
    void *cd;
    H5Eget_auto2(H5E_DEFAULT, &fn, &cd);    /* missing files are errors, not */
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);  /* HDF5 diagnostics on stderr */
    int rc = TP_OK;
    hid_t f = H5Fopen(fname, H5F_ACC_RDONLY, H5P_DEFAULT);
    hid_t g = f >= 0 ? H5Gopen2(f, gname, H5P_DEFAULT) : -1;
    if (g < 0) rc = TP_E_ARG;
    for (int c = 0; c < 6 && rc == TP_OK; c++)
        rc = read_block(g, tp_h5_dsets[c], nx, ny, nz, planes + (size_t)c * nn);
    if (g >= 0) H5Gclose(g);
    if (f >= 0) H5Fclose(f);
    H5Eset_auto2(H5E_DEFAULT, fn, cd);
    return rc;
}

int tp_h5_read_pic(const char *fname, const char *gname, int nx, int ny, int nz, double e0,
                   double b0, float *e, float *b)
{
    if (!e || !b) return TP_E_ARG;
    const size_t nn = (size_t)nx * ny * nz;
    float *planes = (float *)malloc(6 * nn * sizeof(float) + 1);
    if (!planes) return TP_E_NOMEM;
    int rc = tp_h5_read_pic_planes(fname, gname, nx, ny, nz, planes);
    if (rc == TP_OK) {
        /* index_c = i*ny*nz + j*nz + k -> index_f = k*nx*ny + j*nx + i, value
         * scaled in double and stored as float (src/electric_field.c:475-503,
         * src/magnetic_field.c:660-690) */
        const size_t nxy = (size_t)nx * ny, nyz = (size_t)ny * nz;
        for (int c = 0; c < 6; c++) {
            const float *src = planes + (size_t)c * nn;
            float *dst = c < 3 ? e : b;
            const double s = c < 3 ? e0 : b0;
            const int comp = c % 3;
            for (size_t i = 0; i < (size_t)nx; i++)
                for (size_t j = 0; j < (size_t)ny; j++)
                    for (size_t k = 0; k < (size_t)nz; k++)
                        dst[3 * (k * nxy + j * nx + i) + comp] =
                            (float)((double)src[i * nyz + j * nz + k] * s);
        }
    }
    free(planes);
    return rc;
}

/* HDF5 diagnostics off while a function runs: missing files are errors */
typedef struct {
    H5E_auto2_t fn;
    void *cd;
} quiet_t;

static quiet_t quiet_on(void)
{
    quiet_t q;
    H5Eget_auto2(H5E_DEFAULT, &q.fn, &q.cd);
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    return q;
}

static void quiet_off(quiet_t q) { H5Eset_auto2(H5E_DEFAULT, q.fn, q.cd); }

/* the six open datasets and their file spaces, read in x slabs */
typedef struct {
    hid_t d[6], fs[6];
    int ny, nz;
} h5_src;

static int h5_read(void *user, int slice, int comp, long long s0, long long s1, float *dst)
{
    h5_src *h = (h5_src *)user;
    if (slice != 0) return 1;
    hsize_t count[3] = {(hsize_t)(s1 - s0), (hsize_t)h->ny, (hsize_t)h->nz};
    hsize_t offset[3] = {(hsize_t)s0, 0, 0};
    hid_t ms = H5Screate_simple(3, count, NULL);
    int bad = ms < 0 ||
              H5Sselect_hyperslab(h->fs[comp], H5S_SELECT_SET, offset, NULL, count, NULL) < 0 ||
              H5Dread(h->d[comp], H5T_NATIVE_FLOAT, ms, h->fs[comp], H5P_DEFAULT, dst) < 0;
    if (ms >= 0) H5Sclose(ms);
    return bad;
}

int tp_h5_load_pic(tp_ctx *ctx, const char *fname, const char *gname, int nx, int ny, int nz,
                   const double lo[3], const double hi[3], double e0, double b0, int gather_mode)
{
    if (!ctx || !fname || !gname || nx < 1 || ny < 1 || nz < 1) return TP_E_ARG;
    h5_src h;
    for (int c = 0; c < 6; c++) h.d[c] = h.fs[c] = -1;
    h.ny = ny;
    h.nz = nz;
    quiet_t q = quiet_on();
    int rc = TP_OK;
    hid_t f = H5Fopen(fname, H5F_ACC_RDONLY, H5P_DEFAULT);
    hid_t g = f >= 0 ? H5Gopen2(f, gname, H5P_DEFAULT) : -1;
    if (g < 0) rc = TP_E_ARG;
    for (int c = 0; c < 6 && rc == TP_OK; c++) {
        h.d[c] = H5Dopen2(g, tp_h5_dsets[c], H5P_DEFAULT);
        h.fs[c] = h.d[c] >= 0 ? H5Dget_space(h.d[c]) : -1;
        hsize_t dims[3] = {0, 0, 0};
        if (h.fs[c] < 0 || H5Sget_simple_extent_ndims(h.fs[c]) != 3 ||
            H5Sget_simple_extent_dims(h.fs[c], dims, NULL) < 0 || dims[0] < (hsize_t)nx ||
            dims[1] < (hsize_t)ny || dims[2] < (hsize_t)nz)
            rc = TP_E_ARG;
    }
    if (rc == TP_OK)
        rc = tp_gpu_load_field_grid(ctx, nx, ny, nz, lo, hi, TP_PLANES_X_SLOWEST, 0, e0, b0, 0.0,
                                    gather_mode, h5_read, &h, 0);
    for (int c = 0; c < 6; c++) {
        if (h.fs[c] >= 0) H5Sclose(h.fs[c]);
        if (h.d[c] >= 0) H5Dclose(h.d[c]);
    }
    if (g >= 0) H5Gclose(g);
    if (f >= 0) H5Fclose(f);
    quiet_off(q);
    return rc;
}

int tp_h5_write_pic_planes(const char *fname, const char *gname, const long long dims[3],
                           const float *planes)
{
    if (!fname || !gname || !dims || !planes || dims[0] < 1 || dims[1] < 1 || dims[2] < 1)
        return TP_E_ARG;
    const size_t nn = (size_t)dims[0] * dims[1] * dims[2];
    hsize_t hd[3] = {(hsize_t)dims[0], (hsize_t)dims[1], (hsize_t)dims[2]};
    quiet_t q = quiet_on();
    int rc = TP_OK;
    hid_t f = H5Fcreate(fname, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    hid_t g = f >= 0 ? H5Gcreate2(f, gname, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT) : -1;
    hid_t sp = H5Screate_simple(3, hd, NULL);
    if (g < 0 || sp < 0) rc = TP_E_ARG;
    for (int c = 0; c < 6 && rc == TP_OK; c++) {
        hid_t d = H5Dcreate2(g, tp_h5_dsets[c], H5T_NATIVE_FLOAT, sp, H5P_DEFAULT, H5P_DEFAULT,
                             H5P_DEFAULT);
        if (d < 0 ||
            H5Dwrite(d, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, planes + c * nn) < 0)
            rc = TP_E_ARG;
        if (d >= 0) H5Dclose(d);
    }
    if (sp >= 0) H5Sclose(sp);
    if (g >= 0) H5Gclose(g);
    if (f >= 0) H5Fclose(f);
    quiet_off(q);
    return rc;
}
