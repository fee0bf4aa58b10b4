/*
 * tp_pmap_shim.c -- GPU drop-in for the reference's Poincare map (run_type 2).
 *
 * Build this file INTO THE REFERENCE PROGRAM (it includes the reference's
 * headers and reads its globals), compile the reference's src/poincare_map.c
 * with -Dpoincare_map=poincare_map_cpu, and link libtpgpu.so.  It defines
 * poincare_map with the reference's signature (include/poincare_map.h:4,
 * src/poincare_map.c:39):
 *
 *  - the "Poincare map configuration" block is read by the reference's own
 *    config_poincare_map (src/poincare_map.c:357-514);
 *  - each MPI rank takes the reference's share of the start points
 *    (src/poincare_map.c:61-71) and traces them with tp_gpu_poincare_map on
 *    its GPU, in the reference's crossing mode;
 *  - rank 0 gathers the rows and writes data/poincare_map.h5 in the
 *    reference's layout (group /poincare_map, datasets xpos, ypos, zpos of
 *    [npoints][ncross + 1] doubles, src/poincare_map.c:188-246).  The
 *    reference writes collectively through MPI-IO (parallel HDF5); a serial
 *    write of the gathered rows gives the same file.
 *
 * The field is the reference's B grid of systems 3 and 4 (magc_float, or
 * maga/magb with two time slices); like get_magnetic_field
 * (src/emfields.c:84-95) other systems stop with an error.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <mpi.h>
#include <hdf5.h>
#include "constants.h"
#include "domain.h"
#include "magnetic_field.h"
#include "tp_gpu.h"
#include "tp_host.h"

int config_poincare_map(int mpi_rank, char *config_file_name);

/* reference globals (src/poincare_map.c:22-31, src/emfields.c, the grid
 * and field arrays of src/magnetic_field.c:32-39) */
extern struct domain simul_domain;
extern struct grids simul_grid;
extern int system_type, multi_tframe;
extern struct bfields_float *magc_float, *maga_float, *magb_float;
extern int map_for_single_point, face_cross_section, nhorizontal, nvertical, ncross;
extern double fx0, fy0, fz0, min_horizontal, max_horizontal, min_vertical, max_vertical;
extern double step_fieldline, fieldline_length_max;

static tp_ctx *tp_pmap_ctx;

static int tp_pmap_set_field(tp_ctx *c)
{
    if (system_type != 3 && system_type != 4) return -1;
    double lo[3] = {simul_domain.xmin, simul_domain.ymin, simul_domain.zmin};
    double hi[3] = {simul_domain.xmax, simul_domain.ymax, simul_domain.zmax};
    const size_t nn = (size_t)simul_grid.nx * simul_grid.ny * simul_grid.nz;
    /* run_type 2 loads no electric field (src/domain.c:387-394) */
    float *e0 = calloc(nn * 3 + 1, sizeof(float));
    if (!e0) return TP_E_NOMEM;
    int rc;
    if (multi_tframe == 1)
        rc = tp_gpu_set_field_grid2(c, simul_grid.nx, simul_grid.ny, simul_grid.nz, lo, hi,
                                    e0, (const float *)maga_float, e0,
                                    (const float *)magb_float, simul_grid.dt,
                                    TP_GATHER_REF_COMPAT);
    else
        rc = tp_gpu_set_field_grid(c, simul_grid.nx, simul_grid.ny, simul_grid.nz, lo, hi, e0,
                                   (const float *)magc_float, TP_GATHER_REF_COMPAT);
    free(e0);
    return rc;
}

static int tp_pmap_write_h5(int npoints, int ns, const double *x, const double *y,
                            const double *z)
{
    hid_t f = H5Fcreate("data/poincare_map.h5", H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return -1;
    hid_t g = H5Gcreate(f, "/poincare_map", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    hsize_t dims[2] = {(hsize_t)npoints, (hsize_t)ns};
    hid_t sp = H5Screate_simple(2, dims, NULL);
    const char *names[3] = {"xpos", "ypos", "zpos"};
    const double *data[3] = {x, y, z};
    int rc = 0;
    for (int k = 0; k < 3; k++) {
        hid_t d = H5Dcreate(g, names[k], H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT,
                            H5P_DEFAULT);
        if (d < 0 || H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data[k]) < 0)
            rc = -1;
        if (d >= 0) H5Dclose(d);
    }
    H5Sclose(sp);
    H5Gclose(g);
    H5Fclose(f);
    return rc;
}

int poincare_map(int mpi_rank, int mpi_size, char *config_file_name, domain *sdomain)
{
    if (mpi_rank == 0) printf("***** Entering Poincare *****\n");
    memcpy(&simul_domain, sdomain, sizeof(domain));
    config_poincare_map(mpi_rank, config_file_name);

    /* this rank's points: all of them for a single point (rank 0 traces it),
     * otherwise the reference's block split with the remainder on the low
     * ranks */
    const long long total_ll = map_for_single_point ? 1 : (long long)nhorizontal * nvertical;
    if (total_ll < 1 || total_ll * (ncross + 1LL) > 0x7fffffffLL) {
        /* the MPI gather counts are ints */
        fprintf(stderr, "poincare_map: %lld points x %d slots is out of range\n", total_ll,
                ncross + 1);
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    const int total = (int)total_ll;
    int npoints, nshift;
    if (map_for_single_point) {
        npoints = mpi_rank == 0 ? 1 : 0;
        nshift = 0;
    } else {
        npoints = total / mpi_size;
        nshift = npoints * mpi_rank;
        const int res = total % mpi_size;
        if (res != 0) {
            if (mpi_rank < res) {
                npoints++;
                nshift += mpi_rank;
            } else {
                nshift += res;
            }
        }
    }
    const int ns = ncross + 1;
    double p0[3] = {fx0, fy0, fz0};
    double *all = malloc(sizeof(double) * 3 * (size_t)total + 1);
    double *xp = malloc(sizeof(double) * (size_t)ns * (npoints > 0 ? npoints : 1));
    double *yp = malloc(sizeof(double) * (size_t)ns * (npoints > 0 ? npoints : 1));
    double *zp = malloc(sizeof(double) * (size_t)ns * (npoints > 0 ? npoints : 1));
    int *nc = malloc(sizeof(int) * (npoints > 0 ? npoints : 1));
    if (!all || !xp || !yp || !zp || !nc) {
        fprintf(stderr, "poincare_map: out of memory\n");
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    if (tp_host_pmap_starts(face_cross_section, map_for_single_point, p0, min_horizontal,
                            max_horizontal, min_vertical, max_vertical, nhorizontal, nvertical,
                            all) != total) {
        fprintf(stderr, "poincare_map: bad start-point configuration\n");
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    if (npoints > 0) {
        int rc = TP_OK;
        if (!tp_pmap_ctx) rc = tp_gpu_init(0, &tp_pmap_ctx);
        if (rc == TP_OK) {
            rc = tp_pmap_set_field(tp_pmap_ctx);
            if (rc == -1) {
                printf("ERROR: wrong system type.\n");   /* src/emfields.c:93 */
                exit(1);
            }
        }
        if (rc == TP_OK) {
            tp_pmap_params p;
            memset(&p, 0, sizeof(p));
            p.face = face_cross_section;
            p.ncross = ncross;
            p.mode = TP_PMAP_REF_COMPAT;
            p.step = step_fieldline;
            p.length_max = fieldline_length_max;
            p.lo[0] = simul_domain.xmin; p.lo[1] = simul_domain.ymin; p.lo[2] = simul_domain.zmin;
            p.hi[0] = simul_domain.xmax; p.hi[1] = simul_domain.ymax; p.hi[2] = simul_domain.zmax;
            rc = tp_gpu_poincare_map(tp_pmap_ctx, &p, npoints, all + 3 * (size_t)nshift, xp, yp,
                                     zp, nc, NULL);
        }
        if (rc != TP_OK) {
            fprintf(stderr, "poincare_map: GPU path failed: %s\n", tp_gpu_strerror(rc));
            MPI_Abort(MPI_COMM_WORLD, 1);
        }
        fprintf(stderr, "tp_gpu: poincare_map on the GPU (%d field lines)\n", npoints);
        for (int i = 0; i < npoints; i++)
            if (nc[i] < ncross)   /* src/poincare_map.c:179-182 */
                printf("Reached maximum length for field-line tracing before %d crosses\n",
                       nc[i] + 2);
    }

    /* gather the rows on rank 0, in point order */
    double *gx = xp, *gy = yp, *gz = zp;
    if (mpi_size > 1 && !map_for_single_point) {
        int *counts = malloc(sizeof(int) * mpi_size), *displs = malloc(sizeof(int) * mpi_size);
        const int mine = npoints * ns;
        MPI_Gather(&mine, 1, MPI_INT, counts, 1, MPI_INT, 0, MPI_COMM_WORLD);
        if (mpi_rank == 0) {
            displs[0] = 0;
            for (int r = 1; r < mpi_size; r++) displs[r] = displs[r - 1] + counts[r - 1];
            gx = malloc(sizeof(double) * (size_t)total * ns);
            gy = malloc(sizeof(double) * (size_t)total * ns);
            gz = malloc(sizeof(double) * (size_t)total * ns);
        }
        MPI_Gatherv(xp, mine, MPI_DOUBLE, gx, counts, displs, MPI_DOUBLE, 0, MPI_COMM_WORLD);
        MPI_Gatherv(yp, mine, MPI_DOUBLE, gy, counts, displs, MPI_DOUBLE, 0, MPI_COMM_WORLD);
        MPI_Gatherv(zp, mine, MPI_DOUBLE, gz, counts, displs, MPI_DOUBLE, 0, MPI_COMM_WORLD);
        free(counts);
        free(displs);
    }
    if (mpi_rank == 0 && tp_pmap_write_h5(total, ns, gx, gy, gz) != 0) {
        fprintf(stderr, "poincare_map: cannot write data/poincare_map.h5\n");
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    if (gx != xp) { free(gx); free(gy); free(gz); }
    free(all); free(xp); free(yp); free(zp); free(nc);
    if (mpi_rank == 0) printf("***** Leaving Poincare *****\n");
    return 0;
}
