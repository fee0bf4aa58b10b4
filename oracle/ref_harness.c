/*
 * ref_harness.c -- test infrastructure only.
 *
 * A small driver that links the reference tracer's own hot-path objects
 * (compiled from src/tracking.c, src/emfields.c, src/force_free.c,
 * src/interpolation.c, src/{magnetic,electric,velocity}_field.c, src/wlcs.c,
 * src/diagnostics.c of xiaocanli/test_particle, see oracle/Makefile) and calls
 * them on seeded inputs handed in as files.  Its outputs are the golden
 * vectors the project's CPU oracle (oracle/tpref.c) is pinned against.
 *
 * The reference keeps all of its state in file-scope globals that are set by
 * its setup routines; this driver sets the same globals from a "key value"
 * spec file instead of parsing init.dat, because the reference's own
 * test_particle() cannot run system 2 (src/test_particle.c:51 dereferences the
 * WLCS table) and seeds from time(NULL) (src/particle_info.c:318).
 *
 * Modes (all binary files are native-endian):
 *   fields  SPEC PTS OUT     PTS: n*(x,y,z,t) f64 ; OUT: n*(Bx,By,Bz,Ex,Ey,Ez)
 *                            via get_emf (src/emfields.c:32)
 *   wirz    SPEC PTL K OUT   K steps of tracking_wirz + particle_bc
 *                            (src/tracking.c:749, :1172); OUT: final particles
 *   track   SPEC PTL OUTDIR  the fixed-step loop of src/tracking.c:297 driven
 *                            like src/tracking.c:90-190, dumping the raw
 *                            spectrum counts, D sums and remaining counts
 *   traj    SPEC PTL OUTDIR  the same loop with trajectory diagnostics
 *                            (traj_diagnose = 1, src/tracking.c:440-447)
 *   trajh   SPEC PTL OUTDIR  particle_tracking_hybrid with traj_diagnose = 1
 *   gda     SPEC PATH CT OUT the PIC float field readers (.gda files)
 *   h5      SPEC FILE GROUP OUT  the PIC HDF5 readers read_pic_efields_h5 and
 *                            read_pic_bfields_h5 (src/electric_field.c:447-507,
 *                            src/magnetic_field.c:632-692); OUT: the loaded
 *                            efc_float then magc_float ("-": none); prints
 *                            the readers' wall time
 *   hybrid  SPEC PTL [OUT]   particle_tracking_hybrid (src/tracking.c:659) with
 *                            the t=0 spectrum of src/diagnostics.c:50; writes
 *                            data/espectrum.dat and data/diff_coeffs.dat and
 *                            prints the call's wall time; OUT: the particles
 *                            after the call
 *   dfields SPEC PTS OUT     the reference's double gathers
 *                            get_double_bfield_at_point /
 *                            get_double_efield_at_point
 *                            (src/magnetic_field.c:555-591,
 *                            src/electric_field.c:370-406) on the double grid
 *                            of the spec's dgrid; OUT as for fields.  (Its
 *                            get_emf never calls them: systems 3 and 4 are
 *                            set up with float grids, src/domain.c:357-400.)
 *   diff    SPEC PTL0 PTL DT OUT  calc_diff_coeff (src/diagnostics.c:659)
 *   ellint  SPEC KS OUT      ellint (src/wlcs.c:460); KS: n f64 moduli k,
 *                            OUT: n*(E, K)
 *   pmap    SPEC CFG OUT     Poincare map: the "Poincare map configuration"
 *                            block of CFG read by config_poincare_map
 *                            (src/poincare_map.c:357-514), lines traced with
 *                            next_location (:268-320); OUT: xpos, ypos, zpos,
 *                            each npoints*(ncross+1) f64
 *   pmaph   SPEC CFG OUT     poincare_map (src/poincare_map.c:39) as main
 *                            calls it (src/main.c:44-46), in the build with
 *                            the GPU drop-in; reads back
 *                            data/poincare_map.h5, OUT as for pmap
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <mpi.h>
#include <hdf5.h>
#include <omp.h>

/* Layouts of the reference's structs (include/particle_info.h:6-12,
 * include/domain.h:5-24, include/emfields.h:7-10, include/wlcs.h:24-37,
 * include/force_free.h:30-32).  Redeclared here so that this driver needs
 * only the reference's objects, not its headers' macro environment. */
typedef struct { double x, y, z, qm, ux, uy, uz, t, xshift, yshift, zshift; } particles;
typedef struct {
    double xmin_ptl, xmax_ptl, ymin_ptl, ymax_ptl, zmin_ptl, zmax_ptl;
    double xmin, xmax, ymin, ymax, zmin, zmax, tmin, tmax;
} domain;
typedef struct { int nx, ny, nz, nt; double dx, dy, dz, dt; } grids;
typedef struct { double Bx, By, Bz, Ex, Ey, Ez; } emfields;
typedef struct {
    double x_wr, y_wr, z_wr;
    double costheta_wr, cosphi_wr, sintheta_wr, sinphi_wr;
    double cur_wr, t0;
    double x_lp, y_lp, z_lp, r_lp;
    double cosalpha_lp, cosbeta_lp, sinalpha_lp, sinbeta_lp;
    double cur_lp, omega;
} wlcs;
typedef struct { double B0, lambda0; } param_ff;
typedef struct { float x, y, z; } f3;
typedef struct { double x, y, z; } d3;

/* reference entry points */
void set_variables_tracking(grids *sgrid, domain *sdomain, int stype, int mt);
void set_ptl_params_tracking(int pbc, double qm);
void set_variables_emfields(int stype);
void set_variables_ff(grids *sgrid, domain *sdomain);
void set_variabls_interpolation(grids *sgrid, domain *sdomain, int mt);
void set_variables_bfield(grids *sgrid, domain *sdomain, double B0_field, int mt, int dtype);
void set_variables_efield(grids *sgrid, domain *sdomain, double E0_field, int mt, int dtype);
void set_variables_velocity(grids *sgrid, domain *sdomain, double v0_field, int mt, int dtype);
void read_bfields_binary(char *filepath, int ct, int data_type);
void read_efields_binary(char *filepath, int ct, int data_type);
void read_pic_efields_h5(char *fname, char *gname, int data_type);
void read_pic_bfields_h5(char *fname, char *gname, int data_type);
void initialize_bfield(void);
void initialize_efield(void);
void initialize_vfield(void);
void get_emf(double x, double y, double z, double t, emfields *emf);
void tracking_wirz(particles *ptl, double dt);
void particle_bc(particles *ptl, int *iescape);
void calc_diff_coeff(particles *ptl_init, particles *ptl, double dt,
        double *drr, double *dxx, double *dyy, double *dzz, double *dpp,
        double *duu, double *daa, double *dee);
void calc_energy_spectrum(int mpi_rank, int nptl, particles *ptl,
        int nbins, int nt_out, double pmass, int bc_flag);
void particle_tracking_multi_steps_openmp(int nptl, double dt, int tid, int sstep,
        int estep, int nbins, int nt_out, int einterval_t, int traj_diagnose,
        int nsteps_output, double pmass, int *ntraj_accum, particles *ptl_init,
        int *ntraj_diagnostics_points_array, particles *ptl,
        int *nsteps_ptl_tracking, double espect_private[][1],
        double espect_escape_private[][1], particles *ptl_time,
        int *nptl_remain_time, double *drr, double *dxx, double *dyy, double *dzz,
        double *dpp, double *duu, double *daa, double *dee);
void ellint(double k, double *elle, double *ellf);
void particle_tracking_hybrid(int mpi_rank, int nptl, double dt, int nbins,
        int nt_out, int bc_flag, int nsteps_output, double pmass,
        int *ntraj_accum, int tracking_method, int traj_diagnose,
        particles *ptl_init, int *nsteps_ptl_tracking, particles *ptl,
        particles *ptl_time);

/* Poincare map (src/poincare_map.c:18-31): the configuration globals that
 * config_poincare_map sets, and the RK4 step */
int config_poincare_map(int mpi_rank, char *config_file_name);
void next_location(double *fx, double *fy, double *fz, double dstep);
extern int map_for_single_point, face_cross_section, nhorizontal, nvertical, ncross;
extern double fx0, fy0, fz0, min_horizontal, max_horizontal, min_vertical, max_vertical;
extern double step_fieldline, fieldline_length_max;
/* defined only where the GPU drop-in is linked: the reference's own
 * poincare_map needs parallel HDF5 (H5Pset_fapl_mpio) for its output */
int poincare_map(int mpi_rank, int mpi_size, char *config_file_name, domain *sdomain)
    __attribute__((weak));

/* reference globals this driver sets directly */
extern param_ff param;                          /* src/force_free.c:33 */
extern f3 *magc_float;                          /* src/magnetic_field.c:35 */
extern f3 *efc_float;                           /* src/electric_field.c:35 */
extern f3 *maga_float, *magb_float;             /* two time slices */
extern f3 *efa_float, *efb_float;
extern d3 *vc_double;                           /* src/velocity_field.c:34 */
extern f3 *vc_float;                            /* src/velocity_field.c:35 */
extern d3 *magc_double;                         /* src/magnetic_field.c:34 */
extern d3 *efc_double;                          /* src/electric_field.c:34 */
void get_double_bfield_at_point(double x, double y, double z, double t, double *bx,
                                double *by, double *bz);
void get_double_efield_at_point(double x, double y, double z, double t, double *ex,
                                double *ey, double *ez);
extern wlcs *config;                            /* src/wlcs.c:12 */
extern int nwlcs, iBessel;                      /* src/wlcs.c:13 */
extern double normI;                            /* src/wlcs.c:14 */
extern double emin, emax, logemin, logemax, logde; /* src/diagnostics.c:32 */

typedef struct {
    int system, bc, nbins, nt_out, nx, ny, nz, nwlcs, ibessel, nsteps_output, mt, grid_nt;
    int method;                 /* tracking_method of particle_tracking_hybrid */
    double lo[3], hi[3], tmax, qm, dt, pmass, emin, emax, ff_b0, ff_lambda, grid_dt;
    double gda_b0, gda_e0;
    double wlcs_omega, wlcs_normi;
    char grid[512], vgrid[512], wlcs[512], dgrid[512];
} spec_t;

static void die(const char *m) { fprintf(stderr, "ref_harness: %s\n", m); exit(2); }

static void *slurp(const char *path, size_t *n)
{
    FILE *f = fopen(path, "rb");
    if (!f) die(path);
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    void *p = malloc(sz > 0 ? (size_t)sz : 1);
    if (sz > 0 && fread(p, 1, (size_t)sz, f) != (size_t)sz) die("short read");
    fclose(f);
    *n = (size_t)sz;
    return p;
}

static void spit(const char *path, const void *p, size_t n)
{
    FILE *f = fopen(path, "wb");
    if (!f) die(path);
    if (n && fwrite(p, 1, n, f) != n) die("short write");
    fclose(f);
}

static void read_spec(const char *path, spec_t *s)
{
    memset(s, 0, sizeof(*s));
    s->nbins = 102; s->nt_out = 1; s->qm = 1.0; s->pmass = 1.0; s->ff_b0 = 1.0;
    s->ff_lambda = 1.0; s->emin = 1e-3; s->emax = 1e1; s->nsteps_output = 1;
    s->wlcs_normi = 1.0; s->wlcs_omega = 0.001;
    FILE *f = fopen(path, "r");
    if (!f) die(path);
    char key[64], val[512];
    while (fscanf(f, "%63s %511s", key, val) == 2) {
#define I(k) if (!strcmp(key, #k)) { s->k = atoi(val); continue; }
#define D(k) if (!strcmp(key, #k)) { s->k = strtod(val, NULL); continue; }
#define S(k) if (!strcmp(key, #k)) { strcpy(s->k, val); continue; }
        I(system) I(bc) I(nbins) I(nt_out) I(nx) I(ny) I(nz) I(nwlcs) I(ibessel)
        I(nsteps_output) I(mt) I(grid_nt) I(method)
        D(grid_dt) D(gda_b0) D(gda_e0) D(tmax) D(qm) D(dt) D(pmass) D(emin) D(emax) D(ff_b0) D(ff_lambda)
        D(wlcs_omega) D(wlcs_normi)
        S(grid) S(vgrid) S(wlcs) S(dgrid)
        if (!strcmp(key, "xmin")) { s->lo[0] = strtod(val, NULL); continue; }
        if (!strcmp(key, "ymin")) { s->lo[1] = strtod(val, NULL); continue; }
        if (!strcmp(key, "zmin")) { s->lo[2] = strtod(val, NULL); continue; }
        if (!strcmp(key, "xmax")) { s->hi[0] = strtod(val, NULL); continue; }
        if (!strcmp(key, "ymax")) { s->hi[1] = strtod(val, NULL); continue; }
        if (!strcmp(key, "zmax")) { s->hi[2] = strtod(val, NULL); continue; }
        fprintf(stderr, "unknown key %s\n", key);
        die("bad spec");
#undef I
#undef D
#undef S
    }
    fclose(f);
}

/* Mirrors the setup src/domain.c:314-409 performs per system type, with
 * the field arrays filled from files instead of HDF5/binary readers. */
static domain g_dom;

static void setup(const spec_t *s)
{
    domain dom;
    grids g;
    memset(&dom, 0, sizeof(dom));
    memset(&g, 0, sizeof(g));
    dom.xmin = s->lo[0]; dom.ymin = s->lo[1]; dom.zmin = s->lo[2];
    dom.xmax = s->hi[0]; dom.ymax = s->hi[1]; dom.zmax = s->hi[2];
    dom.xmin_ptl = dom.xmin; dom.xmax_ptl = dom.xmax;
    dom.ymin_ptl = dom.ymin; dom.ymax_ptl = dom.ymax;
    dom.zmin_ptl = dom.zmin; dom.zmax_ptl = dom.zmax;
    dom.tmin = 0.0; dom.tmax = s->tmax;
    g_dom = dom;
    if (s->nx > 1) {
        g.nx = s->nx; g.ny = s->ny; g.nz = s->nz; g.nt = s->grid_nt > 0 ? s->grid_nt : 1;
        g.dx = (dom.xmax - dom.xmin) / (g.nx - 1.0);   /* src/domain.c:548 */
        g.dy = (dom.ymax - dom.ymin) / (g.ny - 1.0);
        g.dz = (dom.zmax - dom.zmin) / (g.nz - 1.0);
        g.dt = s->grid_dt;
    }
    const int mt = s->mt;   /* two time slices of the PIC fields */
    set_variables_tracking(&g, &dom, s->system, mt);
    set_ptl_params_tracking(s->bc, s->qm);
    set_variables_emfields(s->system);
    set_variabls_interpolation(&g, &dom, mt);
    size_t nn = (size_t)g.nx * g.ny * g.nz, n;
    if (s->system == 2) {
        set_variables_ff(&g, &dom);
        param.B0 = s->ff_b0;
        param.lambda0 = s->ff_lambda;
        set_variables_velocity(&g, &dom, 1.0, 0, (int)sizeof(double));
        initialize_vfield();
        if (s->vgrid[0]) {
            void *v = slurp(s->vgrid, &n);
            if (n != nn * sizeof(d3)) die("vgrid size");
            memcpy(vc_double, v, n);
            free(v);
        } else {
            memset(vc_double, 0, nn * sizeof(d3));
        }
    } else if (s->system == 4 && s->gda_b0 > 0.0) {
        /* fields come from .gda files (mode "gda") or HDF5 (mode "h5") */
        set_variables_bfield(&g, &dom, s->gda_b0, mt, (int)sizeof(float));
        set_variables_efield(&g, &dom, s->gda_e0, mt, (int)sizeof(float));
        initialize_bfield();
        initialize_efield();
    } else if (s->system == 4) {
        set_variables_bfield(&g, &dom, 1.0, mt, (int)sizeof(float));
        set_variables_efield(&g, &dom, 1.0, mt, (int)sizeof(float));
        initialize_bfield();
        initialize_efield();
        float *v = slurp(s->grid, &n);
        if (!mt) {                          /* E (float3 AoS) then B */
            if (n != 2 * nn * sizeof(f3)) die("grid size");
            memcpy(efc_float, v, nn * sizeof(f3));
            memcpy(magc_float, (char *)v + nn * sizeof(f3), nn * sizeof(f3));
        } else {                            /* E, B of slice a, then of slice b */
            if (n != 4 * nn * sizeof(f3)) die("grid size");
            memcpy(efa_float, v, nn * sizeof(f3));
            memcpy(maga_float, (char *)v + nn * sizeof(f3), nn * sizeof(f3));
            memcpy(efb_float, (char *)v + 2 * nn * sizeof(f3), nn * sizeof(f3));
            memcpy(magb_float, (char *)v + 3 * nn * sizeof(f3), nn * sizeof(f3));
        }
        free(v);
    } else if (s->system == 3) {
        /* MHD fields (src/domain.c:357-371): float velocity and magnetic
         * nodes, one time slice; the file holds v (float3 AoS) then B */
        set_variables_velocity(&g, &dom, 1.0, mt, (int)sizeof(float));
        set_variables_bfield(&g, &dom, 1.0, mt, (int)sizeof(float));
        initialize_vfield();
        initialize_bfield();
        if (mt) die("system 3 with two slices is not tracked by the reference");
        float *v = slurp(s->grid, &n);
        if (n != 2 * nn * sizeof(f3)) die("grid size");
        memcpy(vc_float, v, nn * sizeof(f3));
        memcpy(magc_float, (char *)v + nn * sizeof(f3), nn * sizeof(f3));
        free(v);
    } else if (s->system == 1) {
        /* same table read_wlcs (src/wlcs.c:53-222) builds */
        double *v = slurp(s->wlcs, &n);
        if (n != (size_t)s->nwlcs * 14 * sizeof(double)) die("wlcs size");
        nwlcs = s->nwlcs;
        iBessel = s->ibessel;
        normI = s->wlcs_normi;
        config = malloc(sizeof(wlcs) * nwlcs);
        for (int i = 0; i < nwlcs; i++) {
            const double *r = v + 14 * i;
            config[i].cur_wr = r[0]; config[i].x_wr = r[1];
            config[i].y_wr = r[2]; config[i].z_wr = r[3];
            config[i].costheta_wr = cos(r[4]); config[i].sintheta_wr = sin(r[4]);
            config[i].cosphi_wr = cos(r[5]); config[i].sinphi_wr = sin(r[5]);
            config[i].t0 = r[6]; config[i].cur_lp = r[7]; config[i].r_lp = r[8];
            config[i].x_lp = r[9]; config[i].y_lp = r[10]; config[i].z_lp = r[11];
            config[i].cosalpha_lp = cos(r[12]); config[i].sinalpha_lp = sin(r[12]);
            config[i].cosbeta_lp = cos(r[13]); config[i].sinbeta_lp = sin(r[13]);
            config[i].omega = s->wlcs_omega;
            config[i].cur_wr *= normI;
            config[i].cur_lp *= normI;
        }
        free(v);
    }
    if (s->dgrid[0]) {
        /* double E and B nodes (d3 AoS, E then B) for mode dfields */
        set_variables_bfield(&g, &dom, 1.0, 0, (int)sizeof(double));
        set_variables_efield(&g, &dom, 1.0, 0, (int)sizeof(double));
        initialize_bfield();
        initialize_efield();
        double *v = slurp(s->dgrid, &n);
        if (n != 2 * nn * sizeof(d3)) die("dgrid size");
        memcpy(efc_double, v, nn * sizeof(d3));
        memcpy(magc_double, (char *)v + nn * sizeof(d3), nn * sizeof(d3));
        free(v);
    }
    /* spectrum parameters as src/diagnostics.c:629-631 derives them */
    emin = s->emin;
    emax = s->emax;
    logemin = log10(emin);
    logemax = log10(emax);
    logde = (logemax - logemin) / (s->nbins - 2);
}

int main(int argc, char **argv)
{
    if (argc < 3) die("usage: ref_harness MODE SPEC ...");
    MPI_Init(&argc, &argv);
    spec_t s;
    read_spec(argv[2], &s);
    setup(&s);
    size_t n;
    const char *mode = argv[1];
    if (!strcmp(mode, "fields")) {
        double *pts = slurp(argv[3], &n);
        size_t np = n / (4 * sizeof(double));
        emfields *out = malloc(np * sizeof(emfields) + 1);
        for (size_t i = 0; i < np; i++)
            get_emf(pts[4 * i], pts[4 * i + 1], pts[4 * i + 2], pts[4 * i + 3], &out[i]);
        spit(argv[4], out, np * sizeof(emfields));
    } else if (!strcmp(mode, "dfields")) {
        if (!s.dgrid[0]) die("dfields needs dgrid");
        double *pts = slurp(argv[3], &n);
        size_t np = n / (4 * sizeof(double));
        emfields *out = malloc(np * sizeof(emfields) + 1);
        for (size_t i = 0; i < np; i++) {
            const double *q = pts + 4 * i;
            get_double_bfield_at_point(q[0], q[1], q[2], q[3], &out[i].Bx, &out[i].By,
                                       &out[i].Bz);
            get_double_efield_at_point(q[0], q[1], q[2], q[3], &out[i].Ex, &out[i].Ey,
                                       &out[i].Ez);
        }
        spit(argv[4], out, np * sizeof(emfields));
    } else if (!strcmp(mode, "wirz")) {
        particles *p = slurp(argv[3], &n);
        size_t np = n / sizeof(particles);
        int k = atoi(argv[4]);
        for (size_t i = 0; i < np; i++) {
            int esc = 0;
            for (int j = 0; j < k; j++) {
                tracking_wirz(&p[i], s.dt);
                particle_bc(&p[i], &esc);
            }
        }
        spit(argv[5], p, n);
    } else if (!strcmp(mode, "diff")) {
        particles *p0 = slurp(argv[3], &n);
        size_t np = n / sizeof(particles);
        particles *p = slurp(argv[4], &n);
        double dt = strtod(argv[5], NULL);
        double *out = malloc(np * 8 * sizeof(double) + 1);
        for (size_t i = 0; i < np; i++) {
            double *o = out + 8 * i;
            calc_diff_coeff(&p0[i], &p[i], dt, o, o + 1, o + 2, o + 3, o + 4,
                            o + 5, o + 6, o + 7);
        }
        spit(argv[6], out, np * 8 * sizeof(double));
    } else if (!strcmp(mode, "track")) {
        /* src/tracking.c:90-190 with one thread, keeping the raw arrays */
        particles *p = slurp(argv[3], &n);
        int np = (int)(n / sizeof(particles));
        particles *p0 = malloc(n + 1);
        memcpy(p0, p, n);
        int estep = s.tmax / s.dt;
        int einterval = estep / s.nt_out;
        int nd = estep / 100 + 1;
        int nb = s.nbins, nt = s.nt_out;
        double *dsum = calloc((size_t)8 * nd, sizeof(double));
        /* one spare entry: with open boundaries the reference decrements
         * nptl_remain_time[nd] (src/tracking.c:366-369) */
        int *remain = malloc(sizeof(int) * (nd + 1));
        for (int i = 0; i <= nd; i++) remain[i] = np;
        int *nsteps = malloc(sizeof(int) * (np + 1));
        for (int i = 0; i < np; i++) nsteps[i] = 1;
        double *esp = calloc((size_t)nb * nt + 1, sizeof(double));
        double *esc = calloc((size_t)nb * nt + 1, sizeof(double));
        particle_tracking_multi_steps_openmp(np, s.dt, 0, 1, estep, nb, nt,
            einterval, 0, s.nsteps_output, s.pmass, NULL, p0, NULL, p, nsteps,
            (double (*)[1])esp, (double (*)[1])esc, NULL, remain,
            dsum, dsum + nd, dsum + 2 * nd, dsum + 3 * nd, dsum + 4 * nd,
            dsum + 5 * nd, dsum + 6 * nd, dsum + 7 * nd);
        char path[1024];
        snprintf(path, sizeof(path), "%s/ptl.bin", argv[4]);
        spit(path, p, n);
        snprintf(path, sizeof(path), "%s/nsteps.bin", argv[4]);
        spit(path, nsteps, sizeof(int) * np);
        snprintf(path, sizeof(path), "%s/espect.bin", argv[4]);
        spit(path, esp, sizeof(double) * nb * nt);
        snprintf(path, sizeof(path), "%s/espect_escape.bin", argv[4]);
        spit(path, esc, sizeof(double) * nb * nt);
        snprintf(path, sizeof(path), "%s/dcoeff.bin", argv[4]);
        spit(path, dsum, sizeof(double) * 8 * nd);
        snprintf(path, sizeof(path), "%s/remain.bin", argv[4]);
        spit(path, remain, sizeof(int) * nd);
    } else if (!strcmp(mode, "traj")) {
        /* src/tracking.c:297-462 with traj_diagnose = 1, one thread: every
         * particle gets estep/nsteps_output + 1 trajectory slots, slot 0 its
         * initial state (as src/diagnostics.c:545-552 fills it) */
        particles *p = slurp(argv[3], &n);
        int np = (int)(n / sizeof(particles));
        particles *p0 = malloc(n + 1);
        memcpy(p0, p, n);
        int estep = s.tmax / s.dt;
        int einterval = estep / s.nt_out;
        int nd = estep / 100 + 1;
        int nb = s.nbins, nt = s.nt_out;
        int slots = estep / s.nsteps_output + 1;
        int *accum = malloc(sizeof(int) * (np + 1));
        for (int i = 0; i < np; i++) accum[i] = (i + 1) * slots;
        int *points = malloc(sizeof(int) * (np + 1));
        for (int i = 0; i < np; i++) points[i] = 1;
        particles *traj = calloc((size_t)np * slots + 1, sizeof(particles));
        for (int i = 0; i < np; i++) memcpy(&traj[(size_t)i * slots], &p[i], sizeof(particles));
        double *dsum = calloc((size_t)8 * nd, sizeof(double));
        int *remain = malloc(sizeof(int) * (nd + 1));
        for (int i = 0; i <= nd; i++) remain[i] = np;
        int *nsteps = malloc(sizeof(int) * (np + 1));
        for (int i = 0; i < np; i++) nsteps[i] = 1;
        double *esp = calloc((size_t)nb * nt + 1, sizeof(double));
        double *esc = calloc((size_t)nb * nt + 1, sizeof(double));
        particle_tracking_multi_steps_openmp(np, s.dt, 0, 1, estep, nb, nt,
            einterval, 1, s.nsteps_output, s.pmass, accum, p0, points, p, nsteps,
            (double (*)[1])esp, (double (*)[1])esc, traj, remain,
            dsum, dsum + nd, dsum + 2 * nd, dsum + 3 * nd, dsum + 4 * nd,
            dsum + 5 * nd, dsum + 6 * nd, dsum + 7 * nd);
        char path[1024];
        snprintf(path, sizeof(path), "%s/ptl.bin", argv[4]);
        spit(path, p, n);
        snprintf(path, sizeof(path), "%s/traj.bin", argv[4]);
        spit(path, traj, sizeof(particles) * (size_t)np * slots);
        snprintf(path, sizeof(path), "%s/points.bin", argv[4]);
        spit(path, points, sizeof(int) * np);
        snprintf(path, sizeof(path), "%s/espect.bin", argv[4]);
        spit(path, esp, sizeof(double) * nb * nt);
    } else if (!strcmp(mode, "ellint")) {
        double *ks = slurp(argv[3], &n);
        size_t nk = n / sizeof(double);
        double *out = malloc(nk * 2 * sizeof(double) + 1);
        for (size_t i = 0; i < nk; i++) ellint(ks[i], &out[2 * i], &out[2 * i + 1]);
        spit(argv[4], out, nk * 2 * sizeof(double));
    } else if (!strcmp(mode, "trajh")) {
        /* particle_tracking_hybrid with traj_diagnose = 1 (the call of
         * src/diagnostics.c:555-557) and the spec's tracking_method,
         * estep/nsteps_output + 1 slots per particle, slot 0 the initial
         * state */
        particles *p = slurp(argv[3], &n);
        int np = (int)(n / sizeof(particles));
        particles *p0 = malloc(n + 1);
        memcpy(p0, p, n);
        int estep = s.tmax / s.dt;
        int slots = estep / s.nsteps_output + 1;
        int *accum = malloc(sizeof(int) * (np + 1));
        for (int i = 0; i < np; i++) accum[i] = (i + 1) * slots;
        particles *traj = calloc((size_t)np * slots + 1, sizeof(particles));
        for (int i = 0; i < np; i++) memcpy(&traj[(size_t)i * slots], &p[i], sizeof(particles));
        int *nsteps = malloc(sizeof(int) * (np + 1));
        particle_tracking_hybrid(0, np, s.dt, s.nbins, s.nt_out, s.bc,
            s.nsteps_output, s.pmass, accum, s.method, 1, p0, nsteps, p, traj);
        char path[1024];
        snprintf(path, sizeof(path), "%s/ptl.bin", argv[4]);
        spit(path, p, n);
        snprintf(path, sizeof(path), "%s/traj.bin", argv[4]);
        spit(path, traj, sizeof(particles) * (size_t)np * slots);
    } else if (!strcmp(mode, "gda")) {
        /* read_efields_binary / read_bfields_binary (src/electric_field.c:
         * 257-361, src/magnetic_field.c:440-546) from PATH (with its trailing
         * slash) at time slice CT; writes the loaded float3 arrays */
        int ct = atoi(argv[4]);
        read_efields_binary(argv[3], ct, (int)sizeof(float));
        read_bfields_binary(argv[3], ct, (int)sizeof(float));
        size_t nn = (size_t)s.nx * s.ny * s.nz;
        FILE *fo = fopen(argv[5], "wb");
        if (!fo) die("gda out");
        if (!s.mt) {
            fwrite(efc_float, sizeof(f3), nn, fo);
            fwrite(magc_float, sizeof(f3), nn, fo);
        } else {
            fwrite(efa_float, sizeof(f3), nn, fo);
            fwrite(maga_float, sizeof(f3), nn, fo);
            fwrite(efb_float, sizeof(f3), nn, fo);
            fwrite(magb_float, sizeof(f3), nn, fo);
        }
        fclose(fo);
    } else if (!strcmp(mode, "h5")) {
        /* one time slice (the readers implement multi_tframe = 0 only); prints
         * the readers' wall time; OUT "-" writes nothing */
        if (s.mt) die("h5: one time slice only");
        const double t_start = omp_get_wtime();
        read_pic_efields_h5(argv[3], argv[4], (int)sizeof(float));
        read_pic_bfields_h5(argv[3], argv[4], (int)sizeof(float));
        printf("h5_seconds %.6f\n", omp_get_wtime() - t_start);
        if (strcmp(argv[5], "-") != 0) {
            size_t nn = (size_t)s.nx * s.ny * s.nz;
            FILE *fo = fopen(argv[5], "wb");
            if (!fo) die("h5 out");
            fwrite(efc_float, sizeof(f3), nn, fo);
            fwrite(magc_float, sizeof(f3), nn, fo);
            fclose(fo);
        }
    } else if (!strcmp(mode, "hybrid")) {
        particles *p = slurp(argv[3], &n);
        int np = (int)(n / sizeof(particles));
        particles *p0 = malloc(n + 1);
        memcpy(p0, p, n);
        int *nsteps = malloc(sizeof(int) * (np + 1));
        calc_energy_spectrum(0, np, p, s.nbins, s.nt_out, s.pmass, s.bc);
        const double t_start = omp_get_wtime();
        particle_tracking_hybrid(0, np, s.dt, s.nbins, s.nt_out, s.bc,
            s.nsteps_output, s.pmass, NULL, s.method, 0, p0, nsteps, p, NULL);
        printf("hybrid_seconds %.6f\nhybrid_threads %d\n", omp_get_wtime() - t_start,
               omp_get_max_threads());
        if (argc > 4) spit(argv[4], p, n);
    } else if (!strcmp(mode, "pmap")) {
        /* the tracing loop of src/poincare_map.c:83-185 around the
         * reference's next_location, one MPI rank, lines over OpenMP threads
         * as :122 does; prints the loop's wall time and RK4 step count */
        config_poincare_map(0, argv[3]);
        const int np = map_for_single_point ? 1 : nhorizontal * nvertical;
        const size_t ns = (size_t)ncross + 1;
        double *xp = malloc(sizeof(double) * ns * np), *yp = malloc(sizeof(double) * ns * np),
               *zp = malloc(sizeof(double) * ns * np);
        for (size_t i = 0; i < ns * np; i++) {
            xp[i] = g_dom.xmax * 2;
            yp[i] = g_dom.ymax * 2;
            zp[i] = g_dom.zmax * 2;
        }
        const double dh = nhorizontal > 1 ? (max_horizontal - min_horizontal) / (nhorizontal - 1)
                                          : max_horizontal - min_horizontal;
        const double dv = nvertical > 1 ? (max_vertical - min_vertical) / (nvertical - 1)
                                        : max_vertical - min_vertical;
        long long steps = 0;
        const double t_start = omp_get_wtime();
#pragma omp parallel for reduction(+ : steps)
        for (int ip = 0; ip < np; ip++) {
            double pos[3] = {fx0, fy0, fz0};
            if (!map_for_single_point) {
                int axis = 0;
                for (int a = 0; a < 3; a++) {
                    if (a == face_cross_section) continue;
                    pos[a] = axis == 0 ? dh * (ip % nhorizontal) + min_horizontal
                                       : dv * (ip / nhorizontal) + min_vertical;
                    axis++;
                }
            }
            const double c0 = pos[face_cross_section];
            double *xs = xp + ns * ip, *ys = yp + ns * ip, *zs = zp + ns * ip;
            xs[0] = pos[0]; ys[0] = pos[1]; zs[0] = pos[2];
            double length = 0.0;
            for (int k = 1; k <= ncross; k++) {
                double cross = 1.0;
                while (cross > 0 && length < fieldline_length_max) {
                    double before = pos[face_cross_section];
                    next_location(&pos[0], &pos[1], &pos[2], step_fieldline);
                    cross = (before - c0) * (pos[face_cross_section] - c0);
                    length += step_fieldline;
                    steps++;
                }
                if (!(length < fieldline_length_max)) break;
                xs[k] = pos[0]; ys[k] = pos[1]; zs[k] = pos[2];
            }
        }
        printf("pmap_seconds %.6f\npmap_steps %lld\npmap_threads %d\n",
               omp_get_wtime() - t_start, steps, omp_get_max_threads());
        FILE *fo = fopen(argv[4], "wb");
        if (!fo) die("pmap out");
        fwrite(xp, sizeof(double), ns * np, fo);
        fwrite(yp, sizeof(double), ns * np, fo);
        fwrite(zp, sizeof(double), ns * np, fo);
        fclose(fo);
    } else if (!strcmp(mode, "pmaph")) {
        if (!poincare_map) die("pmaph needs the build with the GPU drop-in");
        poincare_map(0, 1, argv[3], &g_dom);
        hid_t fh = H5Fopen("data/poincare_map.h5", H5F_ACC_RDONLY, H5P_DEFAULT);
        if (fh < 0) die("data/poincare_map.h5");
        FILE *fo = fopen(argv[4], "wb");
        if (!fo) die("pmaph out");
        const char *names[3] = {"/poincare_map/xpos", "/poincare_map/ypos",
                                "/poincare_map/zpos"};
        for (int k = 0; k < 3; k++) {
            hid_t ds = H5Dopen(fh, names[k], H5P_DEFAULT);
            if (ds < 0) die(names[k]);
            hid_t sp = H5Dget_space(ds);
            hsize_t dims[2] = {0, 0};
            if (H5Sget_simple_extent_ndims(sp) != 2) die("pmaph rank");
            H5Sget_simple_extent_dims(sp, dims, NULL);
            size_t cnt = (size_t)dims[0] * dims[1];
            double *buf = malloc(sizeof(double) * cnt + 1);
            if (H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0)
                die("pmaph read");
            fwrite(buf, sizeof(double), cnt, fo);
            free(buf);
            H5Sclose(sp);
            H5Dclose(ds);
        }
        fclose(fo);
        H5Fclose(fh);
    } else {
        die("unknown mode");
    }
    MPI_Finalize();
    return 0;
}
