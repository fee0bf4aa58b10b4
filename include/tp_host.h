/*
 * tp_host.h -- host-side helpers of libtpgpu.so for callers that replace the
 * reference's tracking loop or Poincare map: they turn the raw arrays
 * returned by tp_gpu_track_fixed into the reference's output files, byte for
 * byte, read its field files and lay out its Poincare start points.
 * Plain C; only tp_host_load_pic_gda uses the GPU (through tp_gpu.h).  Each
 * returns 0 or TP_E_ARG (bad arguments or a file that cannot be opened).
 */
#ifndef TP_HOST_H
#define TP_HOST_H

#include "tp_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Bin centres and widths of the energy spectrum, as calc_energy_spectrum
 * and collect_espectrum compute them (src/diagnostics.c:69-78, :227-232);
 * logde = (log10 emax - log10 emin)/(nbins - 2) (src/diagnostics.c:629-631). */
int tp_host_spectrum_bins(int nbins, double emin, double emax, double *ebins,
                          double *einterval);

/* Counts of the t = 0 spectrum with the binning of calc_energy_spectrum
 * (src/diagnostics.c:81-92: truncation, not floor, then clamping). */
int tp_host_spectrum_t0(int n, const tp_particle *ptl, int nbins, double emin,
                        double emax, double *counts);

/* The file calc_energy_spectrum writes (src/diagnostics.c:97-108): bin
 * centres, then counts_t0 / width, "%10.4e " each; with bc_flag 1 also the
 * escape file header with a zero row (src/diagnostics.c:111-122). */
int tp_host_write_espectrum_t0(const char *fname, const char *fname_escape,
                               int nbins, double emin, double emax,
                               const double *counts_t0, int bc_flag);

/* Appends nt_out rows of counts / width (collect_espectrum,
 * src/diagnostics.c:237-247). */
int tp_host_append_espectrum(const char *fname, int nbins, int nt_out,
                             double emin, double emax, const double *counts);

/* data/diff_coeffs.dat (collect_diff_coeffs, src/diagnostics.c:761-771):
 * dcoeff is [8][nd] (drr dxx dyy dzz dpp duu daa dee), nptl_remain [nd]. */
int tp_host_write_diff_coeffs(const char *fname, int nd, const double *dcoeff,
                              const int *nptl_remain);

/* PIC field input from the reference's binary files ("<path>ex_<ct>.gda" ...,
 * nx*ny*nz floats each, x fastest), as read_efields_float_binary and
 * read_bfields_float_binary load them (src/electric_field.c:257-350,
 * src/magnetic_field.c:440-530): values scaled by e0 (= v0 B0) and b0 in
 * double and stored as float, interleaved into the float3 node arrays that
 * tp_gpu_set_field_grid / tp_gpu_set_field_grid2 take.  path includes its
 * trailing separator.
 *   two_slices = 0: slice ct into e_next, b_next;
 *   two_slices = 1: slice min(ct + 1, nt - 1) into e_next, b_next and slice
 *                   max(ct - 1, 0) into e_prev, b_prev (the reference's
 *                   "next" and "previous" slices).
 * Returns TP_E_ARG for a missing or short file (the reference exits). */
int tp_host_read_pic_gda(const char *path, int ct, int nt, int two_slices, int nx, int ny,
                         int nz, double e0, double b0, float *e_next, float *b_next,
                         float *e_prev, float *b_prev);

/* The same files straight onto the GPU: each component file is read in
 * z slabs into pinned staging memory and laid out on the GPU
 * (tp_gpu_load_field_grid, TP_PLANES_X_FASTEST), giving the field
 * tp_gpu_set_field_grid (two_slices = 0) or tp_gpu_set_field_grid2
 * (two_slices = 1, grid_dt) would give for the arrays tp_host_read_pic_gda
 * returns.  TP_E_ARG for a missing or short file. */
int tp_host_load_pic_gda(tp_ctx *ctx, const char *path, int ct, int nt, int two_slices, int nx,
                         int ny, int nz, const double lo[3], const double hi[3], double e0,
                         double b0, double grid_dt, int gather_mode);

/* Start points of a Poincare map (src/poincare_map.c:83-158).  single = 1:
 * the one point p0.  single = 0: nh*nv points on the plane through p0
 * normal to axis face; point ipoint = iv*nh + ih sits at dh*ih + min_h on
 * the first axis other than face and at dv*iv + min_v on the second, with
 * dh = (max_h - min_h)/(nh - 1), or max_h - min_h when nh = 1 (dv alike).
 * starts receives 3 doubles per point.  Returns the number of points, or
 * TP_E_ARG. */
int tp_host_pmap_starts(int face, int single, const double p0[3], double min_h, double max_h,
                        double min_v, double max_v, int nh, int nv, double *starts);

#ifdef __cplusplus
}
#endif
#endif /* TP_HOST_H */
