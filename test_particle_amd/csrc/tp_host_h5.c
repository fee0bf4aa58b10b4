/*
 * tp_host_h5.c -- the reference's PIC HDF5 field input (libtph5.so, the
 * functions of include/tp_h5.h).  Plain C over the serial HDF5 library; the
 * GPU load goes through libtpgpu.so's tp_gpu_load_field_grid.
 */
#include <stddef.h>
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

/* --- trajectory output ------------------------------------------------- */
static const char *const tp_h5_ptl_names[8] = {"x", "y", "z", "qm", "ux", "uy", "uz", "t"};
static const char *const tp_h5_emf_names[6] = {"Bx", "By", "Bz", "Ex", "Ey", "Ez"};

/* compound of k doubles named names[], at offsets 8*j in the file (big
 * endian) or from off[j] in a record of rec bytes in memory (native) */
static hid_t dcompound(int k, const char *const *names, size_t rec, const size_t *off, int file)
{
    hid_t t = H5Tcreate(H5T_COMPOUND, file ? (size_t)8 * k : rec);
    if (t < 0) return -1;
    for (int j = 0; j < k; j++)
        if (H5Tinsert(t, names[j], file ? (size_t)8 * j : off[j],
                      file ? H5T_IEEE_F64BE : H5T_NATIVE_DOUBLE) < 0) {
            H5Tclose(t);
            return -1;
        }
    return t;
}

static const size_t tp_h5_ptl_off[8] = {
    offsetof(tp_particle, x), offsetof(tp_particle, y), offsetof(tp_particle, z),
    offsetof(tp_particle, qm), offsetof(tp_particle, ux), offsetof(tp_particle, uy),
    offsetof(tp_particle, uz), offsetof(tp_particle, t)};
static const size_t tp_h5_seq_off[8] = {0, 8, 16, 24, 32, 40, 48, 56};

static int write_dset(hid_t g, const char *name, int k, const char *const *names, size_t rec,
                      const size_t *off, hsize_t n, const void *buf)
{
    int rc = TP_OK;
    hid_t ft = dcompound(k, names, rec, off, 1), mt = dcompound(k, names, rec, off, 0);
    hid_t sp = H5Screate_simple(1, &n, NULL);
    hid_t d = (ft >= 0 && sp >= 0)
                  ? H5Dcreate2(g, name, ft, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT) : -1;
    if (d < 0 || mt < 0 || (n > 0 && H5Dwrite(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0))
        rc = TP_E_ARG;
    if (d >= 0) H5Dclose(d);
    if (sp >= 0) H5Sclose(sp);
    if (mt >= 0) H5Tclose(mt);
    if (ft >= 0) H5Tclose(ft);
    return rc;
}

int tp_h5_write_particles_fields(const char *fname, int n, const tp_particle *ptl,
                                 const double *emf)
{
    if (!fname || n < 0 || (n > 0 && (!ptl || !emf))) return TP_E_ARG;
    quiet_t q = quiet_on();
    int rc = TP_OK;
    hid_t f = H5Fcreate(fname, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    hid_t g = f >= 0 ? H5Gcreate2(f, "/particles_fields", H5P_DEFAULT, H5P_DEFAULT,
                                  H5P_DEFAULT) : -1;
    if (g < 0) rc = TP_E_ARG;
    if (rc == TP_OK)
        rc = write_dset(g, "particles", 8, tp_h5_ptl_names, sizeof(tp_particle), tp_h5_ptl_off,
                        (hsize_t)n, ptl);
    if (rc == TP_OK)
        rc = write_dset(g, "fields", 6, tp_h5_emf_names, 6 * sizeof(double), tp_h5_seq_off,
                        (hsize_t)n, emf);
    if (g >= 0) H5Gclose(g);
    if (f >= 0 && H5Fclose(f) < 0) rc = TP_E_ARG;
    quiet_off(q);
    return rc;
}

int tp_h5_save_particles_fields(tp_ctx *ctx, const char *fname, int n, const tp_particle *ptl)
{
    if (!ctx || !fname || n < 0 || (n > 0 && !ptl)) return TP_E_ARG;
    double *pts = malloc(sizeof(double) * 4 * (size_t)(n > 0 ? n : 1));
    double *emf = malloc(sizeof(double) * 6 * (size_t)(n > 0 ? n : 1));
    int rc = (pts && emf) ? TP_OK : TP_E_NOMEM;
    for (int i = 0; rc == TP_OK && i < n; i++) {
        pts[4 * i + 0] = ptl[i].x; pts[4 * i + 1] = ptl[i].y;
        pts[4 * i + 2] = ptl[i].z; pts[4 * i + 3] = ptl[i].t;
    }
    if (rc == TP_OK && n > 0) rc = tp_gpu_eval_field(ctx, n, pts, emf, TP_FP64);
    if (rc == TP_OK) rc = tp_h5_write_particles_fields(fname, n, ptl, emf);
    free(pts);
    free(emf);
    return rc;
}

static int read_dset(hid_t g, const char *name, int k, const char *const *names, int cap,
                     int *n, double *buf)
{
    int rc = TP_OK;
    hid_t d = H5Dopen2(g, name, H5P_DEFAULT);
    hid_t sp = d >= 0 ? H5Dget_space(d) : -1;
    hid_t ft = d >= 0 ? H5Dget_type(d) : -1;
    hsize_t dim = 0;
    if (sp < 0 || ft < 0 || H5Sget_simple_extent_ndims(sp) != 1 ||
        H5Tget_class(ft) != H5T_COMPOUND || H5Tget_nmembers(ft) != k) {
        rc = TP_E_ARG;
    } else {
        H5Sget_simple_extent_dims(sp, &dim, NULL);
        if (dim > (hsize_t)cap || (*n >= 0 && (hsize_t)*n != dim)) rc = TP_E_ARG;
    }
    /* the reference's file types: the named members, IEEE doubles (big
     * endian) at offsets 8*j of a record of 8*k bytes */
    if (rc == TP_OK && H5Tget_size(ft) != (size_t)8 * k) rc = TP_E_ARG;
    for (int j = 0; rc == TP_OK && j < k; j++) {
        char *m = H5Tget_member_name(ft, (unsigned)j);
        hid_t mt = H5Tget_member_type(ft, (unsigned)j);
        if (!m || strcmp(m, names[j]) != 0 || mt < 0 || H5Tequal(mt, H5T_IEEE_F64BE) <= 0 ||
            H5Tget_member_offset(ft, (unsigned)j) != (size_t)8 * j)
            rc = TP_E_ARG;
        if (mt >= 0) H5Tclose(mt);
        if (m) H5free_memory(m);
    }
    if (rc == TP_OK) {
        *n = (int)dim;
        hid_t mt = dcompound(k, names, (size_t)8 * k, tp_h5_seq_off, 0);
        if (mt < 0 || (dim > 0 && H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0))
            rc = TP_E_ARG;
        if (mt >= 0) H5Tclose(mt);
    }
    if (ft >= 0) H5Tclose(ft);
    if (sp >= 0) H5Sclose(sp);
    if (d >= 0) H5Dclose(d);
    return rc;
}

int tp_h5_read_particles_fields(const char *fname, int cap, int *n, double *ptl8, double *emf)
{
    if (!fname || !n || cap < 0 || (cap > 0 && (!ptl8 || !emf))) return TP_E_ARG;
    quiet_t q = quiet_on();
    int rc = TP_OK, cnt = -1;
    hid_t f = H5Fopen(fname, H5F_ACC_RDONLY, H5P_DEFAULT);
    hid_t g = f >= 0 ? H5Gopen2(f, "/particles_fields", H5P_DEFAULT) : -1;
    if (g < 0) rc = TP_E_ARG;
    if (rc == TP_OK) rc = read_dset(g, "particles", 8, tp_h5_ptl_names, cap, &cnt, ptl8);
    if (rc == TP_OK) rc = read_dset(g, "fields", 6, tp_h5_emf_names, cap, &cnt, emf);
    if (rc == TP_OK) *n = cnt;
    if (g >= 0) H5Gclose(g);
    if (f >= 0) H5Fclose(f);
    quiet_off(q);
    return rc;
}
