/*
 * tp_h5.h -- PIC field input from the reference's HDF5 files (libtph5.so, a
 * separate library so that libtpgpu.so does not depend on HDF5).
 *
 * The reference's system-4 HDF5 loaders (read_pic_efields_float_h5,
 * src/electric_field.c:447-507, and read_pic_bfields_float_h5,
 * src/magnetic_field.c:632-692, called from src/domain.c:384-398 with file
 * "<path>/T.<tf>/fields_<tf>.h5" and group "Timestep_<tf>") read the float
 * datasets ex, ey, ez and cbx, cby, cbz of the group (the leading
 * nx x ny x nz block of each, x slowest: index (i*ny + j)*nz + k,
 * src/data_io.c:52-75) and store them x fastest, scaled by E0 = v0*B0 and B0
 * in double and rounded to float.  Only one time slice is read from HDF5
 * (the loaders handle multi_tframe = 0 only).  Each function returns 0,
 * TP_E_ARG (missing file, group or dataset, a dataset smaller than the block,
 * bad arguments) or the error of the GPU call.
 */
#ifndef TP_H5_H
#define TP_H5_H

#include "tp_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The six datasets as stored: planes[c*nn + (i*ny + j)*nz + k] for
 * c = ex ey ez cbx cby cbz, unscaled (nn = nx*ny*nz): the planes
 * tp_gpu_set_field_grid_planes takes with order TP_PLANES_X_SLOWEST. */
int tp_h5_read_pic_planes(const char *fname, const char *gname, int nx, int ny, int nz,
                          float *planes);

/* What the reference's loaders leave in efc_float and magc_float: float3
 * nodes x fastest, e[3*((k*ny + j)*nx + i) + c] = (float)(ex * e0) etc. */
int tp_h5_read_pic(const char *fname, const char *gname, int nx, int ny, int nz, double e0,
                   double b0, float *e, float *b);

/* The file straight onto the GPU: each dataset is read in x slabs
 * (hyperslabs) into pinned staging memory and transposed on the GPU
 * (tp_gpu_load_field_grid, TP_PLANES_X_SLOWEST); the field is the one
 * tp_gpu_set_field_grid sets from tp_h5_read_pic's arrays.  lo/hi: the
 * simulation box (node spacing (hi - lo)/(n - 1)). */
int tp_h5_load_pic(tp_ctx *ctx, const char *fname, const char *gname, int nx, int ny, int nz,
                   const double lo[3], const double hi[3], double e0, double b0,
                   int gather_mode);

/* Writes the six datasets ex ey ez cbx cby cbz of dims[0] x dims[1] x
 * dims[2] floats (x slowest) into group gname of a new file fname, planes as
 * for tp_h5_read_pic_planes.  The reference's inputs are written by the PIC
 * code; this helper (no counterpart in the reference) makes files of that
 * layout for tests and synthetic runs. */
int tp_h5_write_pic_planes(const char *fname, const char *gname, const long long dims[3],
                           const float *planes);

/* Trajectory output, the file trajectory_diagnostics writes
 * (data/particle_diagnostics.h5, save_particles_fields,
 * src/diagnostics.c:290-390): group "/particles_fields" holding the
 * datasets "particles" (x y z qm ux uy uz t, the compound type of
 * src/particle_info.c:452-488) and "fields" (Bx By Bz Ex Ey Ez,
 * src/emfields.c:227-257) of n entries each, members big-endian IEEE
 * doubles at offsets 8*k.  emf[6*i] is the field at ptl[i] in that order.
 * The whole file from one process (the reference writes each rank's
 * hyperslab of the same datasets through MPI-IO). */
int tp_h5_write_particles_fields(const char *fname, int n, const tp_particle *ptl,
                                 const double *emf);
/* The same with the field evaluated on the GPU at each point, as the
 * reference calls get_emf at each (src/diagnostics.c:307-310): the
 * context's field through tp_gpu_eval_field in FP64. */
int tp_h5_save_particles_fields(tp_ctx *ctx, const char *fname, int n, const tp_particle *ptl);
/* Reads such a file back as native doubles: ptl8[8*i] = x y z qm ux uy uz t,
 * emf[6*i]; cap is the capacity in entries, *n the count in the file
 * (TP_E_ARG if it exceeds cap or the layout differs from the one
 * written: member names, big-endian doubles, offsets 8*k). */
int tp_h5_read_particles_fields(const char *fname, int cap, int *n, double *ptl8, double *emf);

#ifdef __cplusplus
}
#endif
#endif /* TP_H5_H */
