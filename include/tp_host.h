/*
 * tp_host.h -- host-side helpers of libtpgpu.so for callers that replace the
 * reference's tracking loop: they turn the raw arrays returned by
 * tp_gpu_track_fixed into the reference's output files, byte for byte.
 * Plain C, no GPU use; each returns 0 or TP_E_ARG (bad arguments or a file
 * that cannot be opened).
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

#ifdef __cplusplus
}
#endif
#endif /* TP_HOST_H */
