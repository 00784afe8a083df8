/*
 * hos_driver.c -- C driver of the HOS solver on one MI355X.
 *
 * The call sequence of the reference driver (Csource/2dpar/main_2d_par.c),
 * with the FFTW-MPI set-up replaced by hos_local_size_2d_transposed and every
 * solver call going through the reference's own boundary (include/hos_ref.h):
 *
 *   get_params(initpars.h5)              main_2d_par.c:91
 *   get_ic_2d(initdata.{id}.h5)          :249
 *   Setup_TimeScheme, rhs_hos_setup      :252-254
 *   fft_2d eta/phi, Dealias both         :258-262
 *   Zvel, ifft_2d W^(M), W^(M-2)         :264-267
 *   data0.{id}.h5 (+ data_extra0)        :271-294
 *   while t < T - eps: sol_update_RK     :298-304
 *     on a dtsave crossing: ifft_2d, data{n}.{id}.h5 (+ extras)   :307-353
 *
 * Formats are the reference's (hdf5_routines_2d_par.c): scalar F64 keys in
 * initpars.h5; /eta0 /phi0 /bat0 [Nx][Ny] in initdata; snapshots hold time,
 * g, Nx, Ny, Lx, Ly as [1][1] datasets and eta, phi, bat as [Nx][Ny]; extras
 * hold Array1 = W^(M), Array2 = W^(M-2) of t = 0 (the reference never
 * refreshes them, main_2d_par.c:264-267).  Serial HDF5 (one process).
 *
 * The reference hard-codes scheme 4 and dt = 0.2 (main_2d_par.c:252,302) and
 * NLevs at compile time; here they are options with those defaults:
 *   hos_driver [-d dir] [-n nlevs] [-s scheme] [-t dt] [-h depth] [-f32]
 *              [-q kx_quirk] [-N max_steps] [-S saveflg]
 * (-S overrides the saveflg read from initpars, e.g. to write the data_extra
 * files from an input pair that sets saveflg = 1.)
 */
#include <hdf5.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "../../include/hos_gpu.h"
#include "../../include/hos_ref.h"

#define EPS64 0.00000001 /* main_2d_par.c:13 */

static double read_key(hid_t f, const char *name, int required, double dflt)
{
    if (!H5Lexists(f, name, H5P_DEFAULT)) {
        if (required) {
            fprintf(stderr, "initpars: missing /%s\n", name);
            exit(EXIT_FAILURE);
        }
        return dflt;
    }
    double v[1];
    hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
    if (d < 0 || H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0) {
        fprintf(stderr, "initpars: cannot read /%s\n", name);
        exit(EXIT_FAILURE);
    }
    H5Dclose(d);
    return v[0];
}

/* get_params, hdf5_routines_2d_par.c:13-101 */
static void get_params(const char *fn)
{
    hid_t f = H5Fopen(fn, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) {
        fprintf(stderr, "cannot open %s\n", fn);
        exit(EXIT_FAILURE);
    }
    Nx = (int)read_key(f, "Nx", 1, 0);
    Ny = (int)read_key(f, "Ny", 1, 0);
    Lx = read_key(f, "Lx", 1, 0);
    Ly = read_key(f, "Ly", 1, 0);
    g = read_key(f, "g", 1, 0);
    runsubid = (int)read_key(f, "runsubid", 1, 0);
    T = read_key(f, "T", 1, 0);
    dtsave = read_key(f, "dtsave", 1, 0);
    saveflg = (int)read_key(f, "saveflg", 1, 0);
    rampflg = (int)read_key(f, "rampflg", 1, 0);
    if (rampflg != 0 && rampflg != 1) {
        printf("Illegal value assigned to rampflag, setting to default value (0).\n");
        rampflg = 0;
    }
    Tramp = read_key(f, "Tramp", 1, 0);
    H5Fclose(f);
    Kx0 = 2 * M_PI / Lx;
    Ky0 = 2 * M_PI / Ly;
}

/* ASCII form of the same parameters (README.md:35 names an ASCII input that
 * the reference snapshot does not contain): "key = value" lines, '#'
 * comments, the initpars.h5 keys plus the optional run settings nlevs,
 * scheme, dt, depth and fp32.  Returns 0 if the file does not exist. */
static int get_params_ascii(const char *fn, int *scheme, double *dt)
{
    FILE *f = fopen(fn, "r");
    if (!f) return 0;
    char line[256];
    int seen = 0;
    while (fgets(line, sizeof line, f)) {
        char *h = strchr(line, '#');
        if (h) *h = 0;
        char key[64];
        double v;
        if (sscanf(line, " %63[A-Za-z0-9_] = %lf", key, &v) != 2) continue;
        seen++;
        if (!strcmp(key, "Nx")) Nx = (int)v;
        else if (!strcmp(key, "Ny")) Ny = (int)v;
        else if (!strcmp(key, "Lx")) Lx = v;
        else if (!strcmp(key, "Ly")) Ly = v;
        else if (!strcmp(key, "g")) g = v;
        else if (!strcmp(key, "runsubid")) runsubid = (int)v;
        else if (!strcmp(key, "T")) T = v;
        else if (!strcmp(key, "dtsave")) dtsave = v;
        else if (!strcmp(key, "saveflg")) saveflg = (int)v;
        else if (!strcmp(key, "rampflg")) rampflg = (int)v;
        else if (!strcmp(key, "Tramp")) Tramp = v;
        else if (!strcmp(key, "nlevs")) hos_nlevs = (int)v;
        else if (!strcmp(key, "scheme")) *scheme = (int)v;
        else if (!strcmp(key, "dt")) *dt = v;
        else if (!strcmp(key, "depth")) hos_depth = v;
        else if (!strcmp(key, "fp32")) hos_precision = v != 0 ? HOS_FP32 : HOS_FP64;
        else {
            fprintf(stderr, "%s: unknown key %s\n", fn, key);
            exit(EXIT_FAILURE);
        }
    }
    fclose(f);
    if (!(Nx > 0 && Ny > 0 && Lx > 0 && Ly > 0 && g > 0 && dtsave > 0)) {
        fprintf(stderr, "%s: Nx, Ny, Lx, Ly, g, dtsave are required\n", fn);
        exit(EXIT_FAILURE);
    }
    if (rampflg != 0 && rampflg != 1) rampflg = 0;
    Kx0 = 2 * M_PI / Lx;
    Ky0 = 2 * M_PI / Ly;
    return seen;
}

/* XDMF index for ParaView (xdmf_routines_2d_par.c:26-146): one temporal
 * collection of Polyvertex grids over data_coordinates.h5:/xy.  Unlike the
 * reference, /xy holds the physical coordinates (x, y) of node i*Ny + j (the
 * reference writes wavenumbers there), the data type is spelled out, and the
 * coordinates are built on the heap rather than in a stack array. */
static void xml_grid(FILE *x, double t, const char *datafile)
{
    const long n = (long)Nx * Ny;
    fprintf(x, "<Grid Name=\"Cells\" GridType=\"Uniform\">\n<Time Value=\"%f\"/>\n", t);
    fprintf(x, "<Topology TopologyType=\"Polyvertex\" NodesPerElement=\"%ld\">\n</Topology>\n", n);
    fprintf(x, "<Geometry GeometryType=\"XY\">\n<DataItem DataType=\"Float\" Precision=\"8\" Dimensions=\"%ld 2\" "
               "Format=\"HDF\">data_coordinates.h5:/xy</DataItem>\n</Geometry>\n", n);
    const char *names[3][2] = {{"Phi", "phi"}, {"Eta", "eta"}, {"Bathy", "bat"}};
    for (int q = 0; q < 3; q++)
        fprintf(x, "<Attribute AttributeType=\"Scalar\" Center=\"Node\" Name=\"%s\">\n<DataItem DataType=\"Float\" "
                   "Precision=\"8\" Dimensions=\"%ld 1\" Format=\"HDF\">%s:/%s</DataItem>\n</Attribute>\n",
                names[q][0], n, datafile, names[q][1]);
    fprintf(x, "</Grid>\n");
}

static FILE *init_xml(double t, const char *datafile)
{
    const long n = (long)Nx * Ny;
    double *xy = malloc(sizeof(double) * 2 * n);
    if (!xy) return NULL;
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++) {
            xy[2 * ((long)i * Ny + j)] = i * Lx / Nx;
            xy[2 * ((long)i * Ny + j) + 1] = j * Ly / Ny;
        }
    hid_t f = H5Fcreate("data_coordinates.h5", H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    hsize_t dim[2] = {(hsize_t)n, 2};
    hid_t sp = H5Screate_simple(2, dim, NULL);
    hid_t d = H5Dcreate2(f, "xy", H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, xy);
    H5Dclose(d);
    H5Sclose(sp);
    H5Fclose(f);
    free(xy);
    FILE *x = fopen("HOS_data.xmf", "w");
    if (!x) return NULL;
    fprintf(x, "<Xdmf>\n<Domain>\n<Grid Name=\"CellTime\" GridType=\"Collection\" CollectionType=\"Temporal\">\n");
    xml_grid(x, t, datafile);
    return x;
}

static void close_xml(FILE *x)
{
    if (!x) return;
    fprintf(x, "</Grid>\n</Domain>\n</Xdmf>\n");
    fclose(x);
}

/* get_ic_2d, hdf5_routines_2d_par.c:104-191: [Nx][Ny] -> padded [Nx][Ny+2] */
static void read_field(hid_t f, const char *name, double *tmp, double *out)
{
    hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
    if (d < 0) {
        fprintf(stderr, "initdata: missing %s\n", name);
        exit(EXIT_FAILURE);
    }
    hid_t sp = H5Dget_space(d);
    hsize_t dims[2] = {0, 0};
    if (H5Sget_simple_extent_ndims(sp) != 2 || H5Sget_simple_extent_dims(sp, dims, NULL) != 2 ||
        dims[0] != (hsize_t)Nx || dims[1] != (hsize_t)Ny) {
        fprintf(stderr, "initdata: %s is not [%d][%d]\n", name, Nx, Ny);
        exit(EXIT_FAILURE);
    }
    if (H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, tmp) < 0) {
        fprintf(stderr, "initdata: cannot read %s\n", name);
        exit(EXIT_FAILURE);
    }
    H5Sclose(sp);
    H5Dclose(d);
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++) out[(size_t)(Ny + 2) * i + j] = tmp[(size_t)Ny * i + j];
}

static void scalar(hid_t f, const char *name, double v)
{
    hsize_t dim[2] = {1, 1};
    hid_t sp = H5Screate_simple(2, dim, NULL);
    hid_t d = H5Dcreate2(f, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v);
    H5Dclose(d);
    H5Sclose(sp);
}

/* write_field_2d / write_extra_2d body: padded -> [Nx][Ny] F64 dataset */
static void field(hid_t f, const char *name, const double *padded, double *tmp)
{
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++) tmp[(size_t)Ny * i + j] = padded[(size_t)(Ny + 2) * i + j];
    hsize_t dim[2] = {(hsize_t)Nx, (hsize_t)Ny};
    hid_t sp = H5Screate_simple(2, dim, NULL);
    hid_t d = H5Dcreate2(f, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (d < 0 || H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, tmp) < 0) {
        fprintf(stderr, "cannot write dataset %s\n", name);
        exit(EXIT_FAILURE);
    }
    H5Dclose(d);
    H5Sclose(sp);
}

/* create_file_2d + write_header_2d (hdf5_routines_2d_par.c:194-288) */
static hid_t snapshot(const char *fn, double t)
{
    hid_t f = H5Fcreate(fn, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) {
        fprintf(stderr, "cannot create %s\n", fn);
        exit(EXIT_FAILURE);
    }
    scalar(f, "time", t);
    scalar(f, "g", g);
    scalar(f, "Nx", (double)Nx);
    scalar(f, "Ny", (double)Ny);
    scalar(f, "Lx", Lx);
    scalar(f, "Ly", Ly);
    return f;
}

int main(int argc, char **argv)
{
    int scheme = 4, nmax = -1, info = 0, save_override = -1;
    double dt = 0.2;
    const char *dir = NULL;
    for (int a = 1; a < argc; a++) {
        if (!strcmp(argv[a], "-d") && a + 1 < argc) dir = argv[++a];
        else if (!strcmp(argv[a], "-info")) info = 1;
        else if (!strcmp(argv[a], "-n") && a + 1 < argc) hos_nlevs = atoi(argv[++a]);
        else if (!strcmp(argv[a], "-s") && a + 1 < argc) scheme = atoi(argv[++a]);
        else if (!strcmp(argv[a], "-t") && a + 1 < argc) dt = atof(argv[++a]);
        else if (!strcmp(argv[a], "-h") && a + 1 < argc) hos_depth = atof(argv[++a]);
        else if (!strcmp(argv[a], "-q") && a + 1 < argc) hos_kx_quirk = atoi(argv[++a]);
        else if (!strcmp(argv[a], "-N") && a + 1 < argc) nmax = atoi(argv[++a]);
        else if (!strcmp(argv[a], "-S") && a + 1 < argc) save_override = atoi(argv[++a]);
        else if (!strcmp(argv[a], "-f32")) hos_precision = HOS_FP32;
        else {
            fprintf(stderr, "usage: %s [-d dir] [-n nlevs] [-s scheme] [-t dt] [-h depth] [-f32] [-q 0|1] [-N steps] [-S saveflg]\n",
                    argv[0]);
            return 2;
        }
    }
    if (dir && chdir(dir) != 0) {
        perror(dir);
        return 1;
    }
    struct timespec w0, w1;
    clock_gettime(CLOCK_MONOTONIC, &w0);

    if (access("initpars.h5", R_OK) == 0)
        get_params("initpars.h5");
    else if (!get_params_ascii("initpars.txt", &scheme, &dt)) {
        fprintf(stderr, "neither initpars.h5 nor initpars.txt found\n");
        return 1;
    }
    if (save_override >= 0) saveflg = save_override;
    char subid[2], name[128];
    snprintf(subid, 2, "%d", runsubid); /* one digit, main_2d_par.c:93 */
    snprintf(name, sizeof name, "initdata.%s.h5", subid);
    printf("Reading initial data from:\t%s\n", name);

    hos_local_size_2d_transposed(Nx, Ny / 2 + 1, NULL, NULL, NULL, NULL);
    const size_t al = (size_t)alloc_local;
    double *eta = calloc(4 * al, sizeof(double)), *phi = eta + 2 * al;
    double *bat = calloc(2 * al, sizeof(double));
    double *velwM = calloc(2 * al, sizeof(double)), *velwM2 = calloc(2 * al, sizeof(double));
    double *tmp = calloc((size_t)Nx * Ny, sizeof(double));
    hos_cplx *heta = calloc(2 * al, sizeof(hos_cplx)), *hphi = heta + al;
    hos_cplx *hw[4];
    for (int q = 0; q < 4; q++) hw[q] = calloc(al, sizeof(hos_cplx));
    if (!eta || !bat || !velwM || !velwM2 || !tmp || !heta || !hw[3]) {
        fprintf(stderr, "out of host memory\n");
        return 1;
    }
    hid_t f = H5Fopen(name, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) {
        fprintf(stderr, "cannot open %s\n", name);
        return 1;
    }
    read_field(f, "/eta0", tmp, eta);
    read_field(f, "/phi0", tmp, phi);
    read_field(f, "/bat0", tmp, bat);
    H5Fclose(f);
    if (info) {
        /* parsed inputs only, no GPU work: one line of key=value */
        double se = 0, see = 0, sp = 0, sb = 0;
        for (int i = 0; i < Nx; i++)
            for (int j = 0; j < Ny; j++) {
                const size_t q = (size_t)(Ny + 2) * i + j;
                se += eta[q];
                see += eta[q] * eta[q];
                sp += phi[q] * phi[q];
                sb += fabs(bat[q]);
            }
        const double n = (double)Nx * Ny;
        printf("Nx=%d Ny=%d Lx=%.17g Ly=%.17g g=%.17g runsubid=%d T=%.17g dtsave=%.17g saveflg=%d rampflg=%d "
               "Tramp=%.17g eta_mean=%.17g eta_std=%.17g phi_rms=%.17g bat_abs=%.17g\n",
               Nx, Ny, Lx, Ly, g, runsubid, T, dtsave, saveflg, rampflg, Tramp, se / n,
               sqrt(see / n - (se / n) * (se / n)), sqrt(sp / n), sb);
        return 0;
    }

    FILE *rt = fopen("runtime.dat", "w+"); /* main_2d_par.c:84-87 */
    if (rt) fclose(rt);
    Setup_TimeScheme(scheme);
    rhs_hos_setup();
    double t = 0.0;
    fft_2d(eta, heta, NULL);
    fft_2d(phi, hphi, NULL);
    Dealias(heta);
    Dealias(hphi);
    Zvel(heta, hw[0], hw[1], hw[2], hw[3], t);
    ifft_2d(hw[0], velwM, NULL);
    ifft_2d(hw[1], velwM2, NULL);

    int nfld = 0;
    snprintf(name, sizeof name, "data%d.%s.h5", nfld, subid);
    hid_t s = snapshot(name, t);
    field(s, "eta", eta, tmp); /* the raw IC, as main_2d_par.c:271-273 */
    field(s, "phi", phi, tmp);
    field(s, "bat", bat, tmp);
    H5Fclose(s);
    FILE *xml = init_xml(t, name); /* main_2d_par.c:276-277 */
    printf("Datafile '%s' written at t=%f\n", name, t);
    if (saveflg > 1) {
        snprintf(name, sizeof name, "data_extra%d.%s.h5", nfld, subid);
        s = snapshot(name, t);
        field(s, "Array1", velwM, tmp);
        field(s, "Array2", velwM2, tmp);
        H5Fclose(s);
        printf("Datafile '%s' written at t=%f\n", name, t);
    }

    long nsteps = 0;
    while (t < T - EPS64 && (nmax < 0 || nsteps < nmax)) {
        const double t_old = t;
        sol_update_RK(heta, &t, dt, NULL);
        nsteps++;
        if (floor(t * (1 + EPS64) / dtsave) > floor(t_old * (1 + EPS64) / dtsave)) {
            nfld++;
            ifft_2d(heta, eta, NULL);
            ifft_2d(hphi, phi, NULL);
            snprintf(name, sizeof name, "data%d.%s.h5", nfld, subid);
            s = snapshot(name, t);
            field(s, "eta", eta, tmp);
            field(s, "phi", phi, tmp);
            field(s, "bat", bat, tmp);
            H5Fclose(s);
            if (xml) xml_grid(xml, t, name); /* write_xml, main_2d_par.c:328 */
            printf("Datafile '%s' written at t=%f\n", name, t);
            if (saveflg > 1) {
                snprintf(name, sizeof name, "data_extra%d.%s.h5", nfld, subid);
                s = snapshot(name, t);
                field(s, "Array1", velwM, tmp);
                field(s, "Array2", velwM2, tmp);
                H5Fclose(s);
                printf("Datafile '%s' written at t=%f\n", name, t);
            }
        }
    }
    close_xml(xml);
    hos_ref_sync();
    rhs_hos_clean();
    clock_gettime(CLOCK_MONOTONIC, &w1);
    printf("Total computational time: %10.2f s (%ld steps, wall clock).\n",
           (w1.tv_sec - w0.tv_sec) + 1e-9 * (w1.tv_nsec - w0.tv_nsec), nsteps);
    free(eta);
    free(bat);
    free(velwM);
    free(velwM2);
    free(tmp);
    free(heta);
    for (int q = 0; q < 4; q++) free(hw[q]);
    return 0;
}
